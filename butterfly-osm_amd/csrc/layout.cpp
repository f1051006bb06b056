// layout.cpp — see layout.hpp.
#include "layout.hpp"

#include <algorithm>
#include <atomic>
#include <cstdlib>
#include <numeric>
#include <thread>
#include <tuple>

namespace bfg {

namespace {
template <class T>
void prefix(std::vector<T> &v);

unsigned worker_count() {
    unsigned nt = std::thread::hardware_concurrency();
    if (const char *e = std::getenv("OMP_NUM_THREADS")) {
        const int v = std::atoi(e);
        if (v > 0) nt = (unsigned)v;
    }
    return std::min(std::max(nt, 1u), 64u);
}

// fn(begin, end) over [0, count) in blocks of `grain`, on worker_count()
// threads taking blocks in turn.  Every block runs exactly once; the calls
// must not depend on each other.
template <class F>
void parallel_blocks(size_t count, size_t grain, F &&fn) {
    const unsigned nt = worker_count();
    grain = std::max<size_t>(grain, 1);
    if (nt == 1 || count <= grain) {
        if (count) fn((size_t)0, count);
        return;
    }
    std::atomic<size_t> next{0};
    auto work = [&] {
        for (;;) {
            const size_t b = next.fetch_add(grain);
            if (b >= count) return;
            fn(b, std::min(count, b + grain));
        }
    };
    std::vector<std::thread> th;
    for (unsigned t = 1; t < nt; ++t) th.emplace_back(work);
    work();
    for (auto &t : th) t.join();
}

struct Arc3 {
    uint32_t key, col, w;  // key: source (or tail), col: column in a 64-wide block
};

// Columns of a row for a block of `cols` columns: 16, 32 or 64.
uint32_t width_class(uint32_t cols) { return cols <= 16 ? 16u : cols <= 32 ? 32u : 64u; }

// Appends the rows of one block of W columns from arcs (key, column, weight):
// one row per distinct key, ascending, holding the key and the weights of its
// arcs to the W columns (INF where there is none; parallel arcs keep the
// minimum).  Returns the number of rows.
uint32_t build_rows(std::vector<Arc3> &arcs, uint32_t W, std::vector<uint32_t> &rkey, std::vector<uint32_t> &rw) {
    std::sort(arcs.begin(), arcs.end(), [](const Arc3 &a, const Arc3 &b) {
        return a.key != b.key ? a.key < b.key : a.col < b.col;
    });
    uint32_t rows = 0;
    for (size_t i = 0; i < arcs.size();) {
        const uint32_t key = arcs[i].key;
        const size_t base = rw.size();
        rw.resize(base + W, INF);
        for (; i < arcs.size() && arcs[i].key == key; ++i)
            rw[base + arcs[i].col] = std::min(rw[base + arcs[i].col], arcs[i].w);
        rkey.push_back(key);
        ++rows;
    }
    return rows;
}

// Appends the DOWN rows of one segment from arcs (source, column, weight):
// one row per distinct source holding its arcs to the first 8k columns
// (k octets: the fewest that cover its last arc, or the widest row's of a
// segment of at most 64 rows; INF where there is none, parallel arcs keep
// the minimum), rows ordered by k descending and source ascending.
// run[c] = rows with at least 8 - c octets.
void build_runs(std::vector<Arc3> &arcs, std::vector<uint32_t> &rkey, std::vector<uint32_t> &rw, uint32_t run[8]) {
    std::sort(arcs.begin(), arcs.end(), [](const Arc3 &a, const Arc3 &b) {
        return a.key != b.key ? a.key < b.key : a.col < b.col;
    });
    std::vector<std::pair<uint32_t, size_t>> rows;  // (octets, first arc)
    for (size_t i = 0; i < arcs.size();) {
        size_t e = i;
        uint32_t last = 0;
        for (; e < arcs.size() && arcs[e].key == arcs[i].key; ++e) last = std::max(last, arcs[e].col);
        rows.emplace_back(last / 8 + 1, i);
        i = e;
    }
    // a segment of at most 64 rows is one unit of the DOWN sweep whatever its
    // row widths: its rows share the widest width, so it stays one run
    if (rows.size() <= 64) {
        uint32_t kmax = 0;
        for (const auto &r : rows) kmax = std::max(kmax, r.first);
        for (auto &r : rows) r.first = kmax;
    }
    std::stable_sort(rows.begin(), rows.end(), [](const auto &a, const auto &b) { return a.first > b.first; });
    std::fill(run, run + 8, 0u);
    for (const auto &[k, i] : rows) {
        const uint32_t key = arcs[i].key, W = 8 * k;
        const size_t base = rw.size();
        rw.resize(base + W, INF);
        for (size_t a = i; a < arcs.size() && arcs[a].key == key; ++a)
            rw[base + arcs[a].col] = std::min(rw[base + arcs[a].col], arcs[a].w);
        rkey.push_back(key);
        for (uint32_t c = 8 - k; c < 8; ++c) ++run[c];
    }
}

template <class T>
void prefix(std::vector<T> &v) {  // exclusive prefix sum in place
    T run = 0;
    for (auto &x : v) { T c = x; x = run; run += c; }
}
}  // namespace

bool build_layout(const CchInput &in, const LayoutParams &lp, Layout &L, std::string &err) {
    const uint32_t n = in.n;
    L = Layout();
    L.n = n;
    if (n == 0) { err = "empty graph"; return false; }
    if (!in.up_off || !in.up_tgt || !in.up_w || !in.dn_off || !in.dn_tgt || !in.dn_w) {
        err = "null CSR array";
        return false;
    }
    if (in.up_off[0] != 0 || in.dn_off[0] != 0) { err = "CSR offsets must start at 0"; return false; }
    for (uint32_t r = 0; r < n; ++r)
        if (in.up_off[r + 1] < in.up_off[r] || in.dn_off[r + 1] < in.dn_off[r]) {
            err = "CSR offsets not monotone";
            return false;
        }
    // ---- validation and INF filtering (bucket_ch.rs:191-424) ----
    std::atomic<uint64_t> a_mu{0}, a_md{0}, a_mui{0}, a_mdi{0};
    std::atomic<int> bad{0};  // 1: UP arc not upward, 2: DOWN arc not downward
    parallel_blocks(n, 4096, [&](size_t b, size_t e) {
        uint64_t mu = 0, md = 0, mui = 0, mdi = 0;
        for (uint32_t r = (uint32_t)b; r < e; ++r) {
            for (uint64_t i = in.up_off[r]; i < in.up_off[r + 1]; ++i) {
                const uint32_t v = in.up_tgt[i];
                if (v >= n || v <= r) bad = 1;
                if (in.up_w[i] != INF) ++mu; else ++mui;
            }
            for (uint64_t i = in.dn_off[r]; i < in.dn_off[r + 1]; ++i) {
                if (in.dn_tgt[i] >= r) bad = 2;
                if (in.dn_w[i] != INF) ++md; else ++mdi;
            }
        }
        a_mu += mu; a_md += md; a_mui += mui; a_mdi += mdi;
    });
    if (bad == 1) { err = "UP arc target not above its tail (not a rank-aligned CCH)"; return false; }
    if (bad == 2) { err = "DOWN arc target not below its tail (not a rank-aligned CCH)"; return false; }
    const uint64_t mu = a_mu, md = a_md;
    L.m_up_inf = a_mui;
    L.m_down_inf = a_mdi;
    if (mu >= 0xFFFFFFFFull || md >= 0xFFFFFFFFull) { err = "more than 2^32-1 finite arcs"; return false; }
    L.m_up = mu;
    L.m_down = md;
    // UP arcs by tail rank, duplicates (tail, target) merged to the minimum
    // weight: each rank's finite arcs sorted and merged in place in a copy
    // laid out like the input, then compacted
    std::vector<uint32_t> uo(n + 1, 0), ut, uw;
    {
        const uint64_t mall = in.up_off[n];
        std::vector<std::pair<uint32_t, uint32_t>> tmp(mall);
        parallel_blocks(n, 4096, [&](size_t b, size_t e) {
            for (uint32_t r = (uint32_t)b; r < e; ++r) {
                const uint64_t a0 = in.up_off[r];
                uint64_t k = a0;
                for (uint64_t i = a0; i < in.up_off[r + 1]; ++i)
                    if (in.up_w[i] != INF) tmp[k++] = {in.up_tgt[i], in.up_w[i]};
                std::sort(tmp.begin() + a0, tmp.begin() + k);
                uint64_t o = a0;
                for (uint64_t i = a0; i < k; ++i)
                    if (i == a0 || tmp[i].first != tmp[i - 1].first) tmp[o++] = tmp[i];
                uo[r] = (uint32_t)(o - a0);
            }
        });
        prefix(uo);
        ut.resize(uo[n]);
        uw.resize(uo[n]);
        parallel_blocks(n, 4096, [&](size_t b, size_t e) {
            for (uint32_t r = (uint32_t)b; r < e; ++r)
                for (uint32_t j = 0; j < uo[r + 1] - uo[r]; ++j) {
                    ut[uo[r] + j] = tmp[in.up_off[r] + j].first;
                    uw[uo[r] + j] = tmp[in.up_off[r] + j].second;
                }
        });
    }
    // ---- elimination tree of the symmetrised finite pattern (Liu) ----
    L.parent.assign(n, NONE);
    {
        std::vector<uint32_t> uin_off(n + 1, 0), uin_src(ut.size());
        for (uint32_t t : ut) uin_off[t]++;
        prefix(uin_off);
        std::vector<uint32_t> cur(uin_off.begin(), uin_off.end() - 1);
        for (uint32_t r = 0; r < n; ++r)
            for (uint32_t k = uo[r]; k < uo[r + 1]; ++k) uin_src[cur[ut[k]]++] = r;
        std::vector<uint32_t> anc(n, NONE);
        auto link = [&](uint32_t i, uint32_t j) {
            uint32_t r = i;
            while (anc[r] != NONE && anc[r] != j) { uint32_t t = anc[r]; anc[r] = j; r = t; }
            if (anc[r] == NONE) { anc[r] = j; L.parent[r] = j; }
        };
        for (uint32_t j = 0; j < n; ++j) {
            for (uint64_t i = in.dn_off[j]; i < in.dn_off[j + 1]; ++i)
                if (in.dn_w[i] != INF) link(in.dn_tgt[i], j);
            for (uint32_t k = uin_off[j]; k < uin_off[j + 1]; ++k) link(uin_src[k], j);
        }
    }
    L.depth.assign(n, 0);
    L.subtree.assign(n, 1);
    L.pre.assign(n, 0);
    for (uint32_t v = 0; v < n; ++v)
        if (L.parent[v] != NONE) L.subtree[L.parent[v]] += L.subtree[v];
    {
        std::vector<uint32_t> next(n, 0);
        uint32_t roots_next = 0;
        for (int64_t vv = (int64_t)n - 1; vv >= 0; --vv) {
            uint32_t v = (uint32_t)vv, p = L.parent[v];
            if (p == NONE) { L.depth[v] = 0; L.pre[v] = roots_next; roots_next += L.subtree[v]; }
            else { L.depth[v] = L.depth[p] + 1; L.pre[v] = next[p]; next[p] += L.subtree[v]; }
            next[v] = L.pre[v] + 1;
            L.max_depth = std::max(L.max_depth, L.depth[v]);
        }
    }
    // ---- regions, chains, segments ----
    const uint32_t BL = std::max<uint32_t>(lp.bottom_limit, 1);
    const uint32_t SEG = std::max<uint32_t>(lp.segment_limit, 1);
    std::vector<uint8_t> top(n, 0);
    std::vector<uint32_t> ntopc(n, 0);
    for (uint32_t v = 0; v < n; ++v) {
        top[v] = L.subtree[v] > BL;
        if (top[v] && L.parent[v] != NONE) ntopc[L.parent[v]]++;
    }
    std::vector<uint32_t> chain_of(n, NONE), idx_in_chain(n, 0);
    std::vector<std::vector<uint32_t>> chains;
    for (int64_t vv = (int64_t)n - 1; vv >= 0; --vv) {
        uint32_t v = (uint32_t)vv, p = L.parent[v];
        if (!top[v]) continue;
        if (p == NONE || ntopc[p] >= 2) {
            chain_of[v] = (uint32_t)chains.size();
            chains.emplace_back();
        } else {
            chain_of[v] = chain_of[p];
        }
        idx_in_chain[v] = (uint32_t)chains[chain_of[v]].size();
        chains[chain_of[v]].push_back(v);
    }
    L.n_chains = (uint32_t)chains.size();
    std::vector<uint32_t> chain_seg0(chains.size());
    std::vector<uint32_t> seg_level;
    std::vector<std::pair<uint32_t, uint32_t>> seg_range;  // (chain, first index)
    for (uint32_t c = 0; c < chains.size(); ++c) {
        uint32_t head = chains[c][0], p = L.parent[head];
        uint32_t lvl = 0;
        if (p != NONE) lvl = seg_level[chain_seg0[chain_of[p]] + idx_in_chain[p] / SEG] + 1;
        chain_seg0[c] = (uint32_t)seg_level.size();
        for (uint32_t i = 0; i < chains[c].size(); i += SEG) {
            seg_level.push_back(lvl++);
            seg_range.emplace_back(c, i);
        }
    }
    // bottom heights
    std::vector<uint32_t> height(n, 0);
    for (uint32_t v = 0; v < n; ++v) {
        uint32_t p = L.parent[v];
        if (!top[v] && p != NONE && !top[p]) height[p] = std::max(height[p], height[v] + 1);
    }
    for (uint32_t v = 0; v < n; ++v)
        if (!top[v]) L.max_height = std::max(L.max_height, height[v]);
    // ---- segment order: the segments of one level in pre-order of their heads
    // (neighbours in the tree side by side) ----
    uint32_t n_seglvl = 0;
    for (uint32_t l : seg_level) n_seglvl = std::max(n_seglvl, l + 1);
    std::vector<std::vector<uint32_t>> lvl_segs(n_seglvl);
    for (uint32_t s = 0; s < seg_level.size(); ++s) lvl_segs[seg_level[s]].push_back(s);
    auto seg_len = [&](uint32_t s) {
        const auto &ch = chains[seg_range[s].first];
        return std::min<uint32_t>((uint32_t)ch.size() - seg_range[s].second, SEG);
    };
    std::vector<uint32_t> seg_tile(seg_level.size());  // segment index in position order
    uint32_t n_tiles = 0;
    for (auto &segs : lvl_segs) {
        std::sort(segs.begin(), segs.end(), [&](uint32_t a, uint32_t b) {
            return L.pre[chains[seg_range[a].first][seg_range[a].second]] <
                   L.pre[chains[seg_range[b].first][seg_range[b].second]];
        });
        for (uint32_t s : segs) seg_tile[s] = n_tiles++;
    }
    std::vector<uint32_t> seg_of(n, NONE);  // segment of a top node, in position order
    for (uint32_t v = 0; v < n; ++v)
        if (top[v]) seg_of[v] = seg_tile[chain_seg0[chain_of[v]] + idx_in_chain[v] / SEG];
    // ---- positions ----
    L.pos_of_rank.assign(n, NONE);
    L.rank_of_pos.assign(n, NONE);
    uint32_t pos = 0;
    std::vector<uint32_t> seg_start(n_tiles);  // first position of a segment
    for (uint32_t l = 0; l < n_seglvl; ++l) {
        L.seglvl_seg.push_back((uint32_t)L.seg_pos.size());
        uint32_t cur = NONE;
        for (uint32_t s : lvl_segs[l]) {
            if (seg_tile[s] != cur) {
                cur = seg_tile[s];
                L.seg_pos.push_back(pos);
                seg_start[cur] = pos;
            }
            const auto &ch = chains[seg_range[s].first];
            const uint32_t i0 = seg_range[s].second, i1 = i0 + seg_len(s);
            for (uint32_t i = i0; i < i1; ++i) { L.pos_of_rank[ch[i]] = pos; L.rank_of_pos[pos] = ch[i]; ++pos; }
            L.max_segment = std::max(L.max_segment, pos - seg_start[cur]);
        }
    }
    L.seglvl_seg.push_back((uint32_t)L.seg_pos.size());
    L.seg_pos.push_back(pos);
    L.n_top = pos;
    {
        std::vector<std::vector<uint32_t>> by_h(L.max_height + 1);
        for (uint32_t v = 0; v < n; ++v)
            if (!top[v]) by_h[height[v]].push_back(v);
        for (int64_t h = L.max_height; h >= 0; --h) {
            auto &vs = by_h[h];
            if (vs.empty()) continue;
            std::sort(vs.begin(), vs.end(), [&](uint32_t a, uint32_t b) { return L.pre[a] < L.pre[b]; });
            L.botlvl_pos.push_back(pos);
            for (uint32_t v : vs) { L.pos_of_rank[v] = pos; L.rank_of_pos[pos] = v; ++pos; }
        }
        L.botlvl_pos.push_back(pos);
    }
    if (pos != n) { err = "layout lost nodes"; return false; }
    // ---- per position ----
    L.pos_parent.resize(n); L.pos_pre.resize(n); L.pos_subtree.resize(n); L.pos_depth.resize(n);
    for (uint32_t p = 0; p < n; ++p) {
        uint32_t v = L.rank_of_pos[p];
        L.pos_parent[p] = L.parent[v] == NONE ? NONE : L.pos_of_rank[L.parent[v]];
        L.pos_pre[p] = L.pre[v];
        L.pos_subtree[p] = L.subtree[v];
        L.pos_depth[p] = L.depth[v];
    }
    auto same_seg = [&](uint32_t a, uint32_t b) { return top[a] && top[b] && seg_of[a] == seg_of[b]; };
    // ---- DOWN pull by head position, outside arcs first ----
    {
        std::vector<uint32_t> cnt(n + 1, 0);
        parallel_blocks(n, 4096, [&](size_t b, size_t e) {
            for (uint32_t r = (uint32_t)b; r < e; ++r)
                for (uint64_t i = in.dn_off[r]; i < in.dn_off[r + 1]; ++i)
                    if (in.dn_w[i] != INF) __atomic_fetch_add(&cnt[L.pos_of_rank[in.dn_tgt[i]]], 1u, __ATOMIC_RELAXED);
        });
        L.in_off = cnt;
        prefix(L.in_off);
        L.in_src.resize(md);
        L.in_w.resize(md);
        // fill in any order; the sort below fixes each head's list
        std::vector<uint32_t> cur(L.in_off.begin(), L.in_off.end() - 1);
        parallel_blocks(n, 4096, [&](size_t b, size_t e) {
            for (uint32_t r = (uint32_t)b; r < e; ++r) {
                const uint32_t pr = L.pos_of_rank[r];
                for (uint64_t i = in.dn_off[r]; i < in.dn_off[r + 1]; ++i) {
                    if (in.dn_w[i] == INF) continue;
                    const uint32_t k = __atomic_fetch_add(&cur[L.pos_of_rank[in.dn_tgt[i]]], 1u, __ATOMIC_RELAXED);
                    L.in_src[k] = pr;
                    L.in_w[k] = in.dn_w[i];
                }
            }
        });
        L.in_split.resize(n);
        std::atomic<uint64_t> intra{0};
        parallel_blocks(n, 1024, [&](size_t pb, size_t pe) {
            std::vector<std::pair<uint32_t, uint32_t>> tmp;
            uint64_t my_intra = 0;
            for (uint32_t p = (uint32_t)pb; p < pe; ++p) {
                const uint32_t a = L.in_off[p], b = L.in_off[p + 1];
                const uint32_t v = L.rank_of_pos[p];
                // second class: top head — arcs from its own segment; bottom
                // head — arcs from top sources
                auto second = [&](uint32_t src_pos) {
                    return top[v] ? same_seg(L.rank_of_pos[src_pos], v) : src_pos < L.n_top;
                };
                tmp.resize(b - a);
                for (uint32_t k = a; k < b; ++k) tmp[k - a] = {L.in_src[k], L.in_w[k]};
                std::sort(tmp.begin(), tmp.end(), [&](const std::pair<uint32_t, uint32_t> &x,
                                                      const std::pair<uint32_t, uint32_t> &y) {
                    const bool ix = second(x.first), iy = second(y.first);
                    if (ix != iy) return !ix;
                    return x < y;
                });
                uint32_t split = b;
                for (uint32_t k = a; k < b; ++k) {
                    L.in_src[k] = tmp[k - a].first;
                    L.in_w[k] = tmp[k - a].second;
                    if (split == b && second(tmp[k - a].first)) split = k;
                }
                L.in_split[p] = split;
                if (top[v]) my_intra += b - split;
            }
            intra += my_intra;
        });
        L.m_down_intra = intra;
    }
    // ---- UP by tail position, outside arcs first ----
    {
        L.up_off.assign(n + 1, 0);
        for (uint32_t r = 0; r < n; ++r) L.up_off[L.pos_of_rank[r]] = uo[r + 1] - uo[r];
        prefix(L.up_off);
        L.up_tgt.resize(ut.size());
        L.up_w.resize(ut.size());
        L.up_split.resize(n);
        std::atomic<uint64_t> intra{0};
        parallel_blocks(n, 1024, [&](size_t pb, size_t pe) {
            std::vector<std::pair<uint32_t, uint32_t>> outs, ins;
            uint64_t my_intra = 0;
            for (uint32_t p = (uint32_t)pb; p < pe; ++p) {
                const uint32_t x = L.rank_of_pos[p];
                outs.clear();
                ins.clear();
                for (uint32_t k = uo[x]; k < uo[x + 1]; ++k) {
                    const uint32_t y = ut[k];
                    if (same_seg(x, y)) ins.emplace_back(L.pos_of_rank[y] - seg_start[seg_of[y]], uw[k]);
                    else outs.emplace_back(L.pos_of_rank[y], uw[k]);
                }
                std::sort(outs.begin(), outs.end());
                std::sort(ins.begin(), ins.end());
                uint32_t k = L.up_off[p];
                for (auto &e : outs) { L.up_tgt[k] = e.first; L.up_w[k] = e.second; ++k; }
                L.up_split[p] = k;
                for (auto &e : ins) { L.up_tgt[k] = e.first; L.up_w[k] = e.second; ++k; }
                my_intra += ins.size();
            }
            intra += my_intra;
        });
        L.m_up_intra = intra;
    }
    // ---- depth gaps of the bottom-sourced arcs of bottom heads ----
    {
        const uint32_t ib = L.in_off[L.n_top];
        L.in_dd.assign(L.in_off[n] - ib, 0);
        bool deep = false;
        parallel_blocks(n - L.n_top, 4096, [&](size_t b, size_t e) {
            for (uint32_t p = L.n_top + (uint32_t)b; p < L.n_top + e; ++p)
                for (uint32_t a = L.in_off[p]; a < L.in_split[p]; ++a) {
                    const uint32_t dd = L.pos_depth[p] - L.pos_depth[L.in_src[a]];
                    if (dd > 0xFFFF) deep = true;
                    L.in_dd[a - ib] = (uint16_t)dd;
                }
        });
        if (deep) { err = "bottom subtree deeper than 65535"; return false; }
    }
    // ---- bottom chunks: bottom nodes in pre-order, their top-sourced DOWN arcs
    // grouped by source ----
    const uint32_t BB = std::max<uint32_t>(lp.block_limit, 1);
    {
        std::vector<uint32_t> bot;
        for (uint32_t p = L.n_top; p < n; ++p) bot.push_back(p);
        std::sort(bot.begin(), bot.end(), [&](uint32_t a, uint32_t b) { return L.pos_pre[a] < L.pos_pre[b]; });
        const size_t nch = (bot.size() + BB - 1) / BB;
        struct Chunk {
            std::vector<uint32_t> src, aoff, aw;
            std::vector<uint16_t> av;
        };
        std::vector<Chunk> ch(nch);
        parallel_blocks(nch, 64, [&](size_t cb, size_t ce) {
            std::vector<std::tuple<uint32_t, uint32_t, uint32_t>> arcs;
            for (size_t c = cb; c < ce; ++c) {
                const size_t i = c * BB, e = std::min(bot.size(), i + BB);
                arcs.clear();
                for (size_t k = i; k < e; ++k)
                    for (uint32_t a = L.in_split[bot[k]]; a < L.in_off[bot[k] + 1]; ++a)
                        arcs.emplace_back(L.in_src[a], (uint32_t)(k - i), L.in_w[a]);
                std::sort(arcs.begin(), arcs.end());
                Chunk &q = ch[c];
                for (size_t j = 0; j < arcs.size(); ++j) {
                    if (j == 0 || std::get<0>(arcs[j]) != std::get<0>(arcs[j - 1])) {
                        q.src.push_back(std::get<0>(arcs[j]));
                        q.aoff.push_back((uint32_t)q.av.size());
                    }
                    q.av.push_back((uint16_t)std::get<1>(arcs[j]));
                    q.aw.push_back(std::get<2>(arcs[j]));
                }
            }
        });
        L.ch_pos = bot;
        L.ch_off.clear();
        for (size_t c = 0; c < nch; ++c) {
            L.ch_off.push_back((uint32_t)(c * BB));
            L.cs_off.push_back((uint32_t)L.cs_src.size());
            const uint32_t base = (uint32_t)L.ca_v.size();
            for (size_t j = 0; j < ch[c].src.size(); ++j) {
                L.cs_src.push_back(ch[c].src[j]);
                L.cs_aoff.push_back(base + ch[c].aoff[j]);
            }
            L.ca_v.insert(L.ca_v.end(), ch[c].av.begin(), ch[c].av.end());
            L.ca_w.insert(L.ca_w.end(), ch[c].aw.begin(), ch[c].aw.end());
            Chunk().src.swap(ch[c].src);
        }
        L.ch_off.push_back((uint32_t)L.ch_pos.size());
        L.cs_off.push_back((uint32_t)L.cs_src.size());
        L.cs_aoff.push_back((uint32_t)L.ca_v.size());
    }
    // ---- dense tiles of the top region (segments of at most 64 nodes), built
    // per segment in parallel and concatenated in segment order ----
    {
        const uint32_t nseg = (uint32_t)L.seg_pos.size() - 1;
        if (L.max_segment > 64) { err = "segment_limit above 64"; return false; }
        struct Tiles {
            std::vector<uint32_t> xr_src, xr_w;                 // DOWN rows, runs by width
            uint32_t run[8];
            std::vector<std::pair<uint32_t, uint32_t>> blocks;  // (tab 0/1 * 64 + 8a + o, local block)
            std::vector<uint32_t> tb_w;                         // 64 weights per local block
            std::vector<uint32_t> pt_tgt, pr_cnt, pr_wid, pr_x, pr_w;  // UP push: targets, rows and width per block
        };
        std::vector<Tiles> tl(nseg);
        parallel_blocks(nseg, 16, [&](size_t sb, size_t se) {
            std::vector<Arc3> a3;
            std::vector<std::tuple<uint32_t, uint32_t, uint32_t>> outs;  // (target, tail position, weight)
            std::vector<uint32_t> colv;                                   // target column of each arc of outs
            for (uint32_t s = (uint32_t)sb; s < se; ++s) {
                Tiles &t = tl[s];
                const uint32_t p0 = L.seg_pos[s], len = L.seg_pos[s + 1] - p0;
                a3.clear();
                for (uint32_t j = 0; j < len; ++j)
                    for (uint32_t k = L.in_off[p0 + j]; k < L.in_split[p0 + j]; ++k)
                        a3.push_back({L.in_src[k], j, L.in_w[k]});
                build_runs(a3, t.xr_src, t.xr_w, t.run);
                uint32_t tab[2][64];
                std::fill(&tab[0][0], &tab[0][0] + 128, NONE);
                auto put = [&](int d, uint32_t a, uint32_t o, uint32_t i, uint32_t j, uint32_t w) {
                    uint32_t &k = tab[d][8 * a + o];
                    if (k == NONE) {
                        k = (uint32_t)t.blocks.size();
                        t.blocks.emplace_back(64 * d + 8 * a + o, k);
                        t.tb_w.resize(t.tb_w.size() + 64, INF);
                    }
                    uint32_t &slot = t.tb_w[64 * (size_t)k + 8 * i + j];
                    slot = std::min(slot, w);
                };
                for (uint32_t i = 0; i < len; ++i) {
                    const uint32_t p = p0 + i;
                    for (uint32_t k = L.in_split[p]; k < L.in_off[p + 1]; ++k) {
                        const uint32_t ti = L.in_src[k] - p0;  // tail before head: ti < i
                        put(0, ti / 8, i / 8, ti % 8, i % 8, L.in_w[k]);
                    }
                    for (uint32_t k = L.up_split[p]; k < L.up_off[p + 1]; ++k) {
                        const uint32_t h = L.up_tgt[k];  // head offset, h < i
                        put(1, i / 8, h / 8, i % 8, h % 8, L.up_w[k]);
                    }
                }
                outs.clear();
                for (uint32_t i = 0; i < len; ++i)
                    for (uint32_t k = L.up_off[p0 + i]; k < L.up_split[p0 + i]; ++k)
                        outs.emplace_back(L.up_tgt[k], p0 + i, L.up_w[k]);
                std::sort(outs.begin(), outs.end());
                for (size_t q = 0; q < outs.size(); ++q)
                    if (q == 0 || std::get<0>(outs[q]) != std::get<0>(outs[q - 1]))
                        t.pt_tgt.push_back(std::get<0>(outs[q]));
                const uint32_t nt = (uint32_t)t.pt_tgt.size();
                t.pt_tgt.resize((size_t)(nt + 63) / 64 * 64, NONE);
                colv.resize(outs.size());
                for (size_t q = 0, c = 0; q < outs.size(); ++q) {
                    if (q > 0 && std::get<0>(outs[q]) != std::get<0>(outs[q - 1])) ++c;
                    colv[q] = (uint32_t)c;
                }
                size_t q = 0;
                for (uint32_t kb = 0; kb * 64 < nt; ++kb) {
                    a3.clear();
                    for (; q < outs.size() && colv[q] < 64 * (kb + 1); ++q)
                        a3.push_back({std::get<1>(outs[q]), colv[q] - 64 * kb, std::get<2>(outs[q])});
                    const uint32_t W = width_class(std::min<uint32_t>(64, nt - 64 * kb));
                    t.pr_wid.push_back(W);
                    t.pr_cnt.push_back(build_rows(a3, W, t.pr_x, t.pr_w));
                }
            }
        });
        L.xr_off.assign(nseg + 1, 0);
        L.xr_run.assign(8 * (size_t)nseg, 0);
        L.xw_off.assign(nseg + 1, 0);
        L.pr_wid.clear();
        L.pw_off.clear();
        L.tb_dn.assign(64 * (size_t)nseg, NONE);
        L.tb_up.assign(64 * (size_t)nseg, NONE);
        L.pt_blk.assign(nseg + 1, 0);
        L.xr_src.clear(); L.xr_w.clear(); L.tb_w.clear();
        L.pt_tgt.clear(); L.pr_off.clear(); L.pr_x.clear(); L.pr_w.clear();
        size_t nx = 0, nxw = 0, ntb = 0, npt = 0, npr = 0, nprw = 0;
        for (const Tiles &t : tl) {
            nx += t.xr_src.size(); nxw += t.xr_w.size(); ntb += t.tb_w.size();
            npt += t.pt_tgt.size(); npr += t.pr_x.size(); nprw += t.pr_w.size();
        }
        if (nxw >= (1ull << 32) || nprw >= (1ull << 32) || ntb >= (1ull << 32)) {
            err = "dense tiles exceed 2^32 weights";
            return false;
        }
        L.xr_src.reserve(nx); L.xr_w.reserve(nxw); L.tb_w.reserve(ntb);
        L.pt_tgt.reserve(npt); L.pr_x.reserve(npr); L.pr_w.reserve(nprw);
        for (uint32_t s = 0; s < nseg; ++s) {
            Tiles &t = tl[s];
            L.xr_off[s] = (uint32_t)L.xr_src.size();
            std::copy(t.run, t.run + 8, L.xr_run.begin() + 8 * (size_t)s);
            L.xw_off[s] = (uint32_t)L.xr_w.size();
            L.xr_src.insert(L.xr_src.end(), t.xr_src.begin(), t.xr_src.end());
            L.xr_w.insert(L.xr_w.end(), t.xr_w.begin(), t.xr_w.end());
            const uint32_t b0 = (uint32_t)(L.tb_w.size() / 64);
            for (auto &bk : t.blocks)
                (bk.first < 64 ? L.tb_dn : L.tb_up)[64 * (size_t)s + bk.first % 64] = b0 + bk.second;
            L.tb_w.insert(L.tb_w.end(), t.tb_w.begin(), t.tb_w.end());
            L.pt_blk[s] = (uint32_t)(L.pt_tgt.size() / 64);
            L.pt_tgt.insert(L.pt_tgt.end(), t.pt_tgt.begin(), t.pt_tgt.end());
            uint32_t r = (uint32_t)L.pr_x.size(), w = (uint32_t)L.pr_w.size();
            for (size_t k = 0; k < t.pr_cnt.size(); ++k) {
                L.pr_off.push_back(r);
                L.pw_off.push_back(w);
                L.pr_wid.push_back(t.pr_wid[k]);
                r += t.pr_cnt[k];
                w += t.pr_cnt[k] * t.pr_wid[k];
            }
            L.pr_x.insert(L.pr_x.end(), t.pr_x.begin(), t.pr_x.end());
            L.pr_w.insert(L.pr_w.end(), t.pr_w.begin(), t.pr_w.end());
            Tiles().xr_w.swap(t.xr_w);
            Tiles().pr_w.swap(t.pr_w);
        }
        L.xr_off[nseg] = (uint32_t)L.xr_src.size();
        L.xw_off[nseg] = (uint32_t)L.xr_w.size();
        L.pt_blk[nseg] = (uint32_t)(L.pt_tgt.size() / 64);
        L.pr_off.push_back((uint32_t)L.pr_x.size());
        L.pw_off.push_back((uint32_t)L.pr_w.size());
        L.pk_seg.assign(L.pt_blk[nseg], 0);
        for (uint32_t s = 0; s < nseg; ++s)
            for (uint32_t k = L.pt_blk[s]; k < L.pt_blk[s + 1]; ++k) L.pk_seg[k] = s;
        L.ell_down_entries = L.xr_w.size();
        L.ell_up_entries = L.pr_w.size();
    }
    return true;
}

bool check_layout(const CchInput &in, const Layout &L, std::string &err) {
    const uint32_t n = in.n;
    // the checks run in parallel blocks; the first failure's message is kept
    std::atomic<const char *> first{nullptr};
    auto flag = [&](const char *m) {
        const char *z = nullptr;
        first.compare_exchange_strong(z, m);
    };
    auto failed = [&] {
        if (!first.load()) return false;
        err = first.load();
        return true;
    };
    auto is_anc = [&](uint32_t a, uint32_t d) {  // rank a ancestor-or-self of rank d
        return L.pre[a] <= L.pre[d] && L.pre[d] < L.pre[a] + L.subtree[a];
    };
    parallel_blocks(n, 4096, [&](size_t b, size_t e) {
        for (uint32_t r = (uint32_t)b; r < e; ++r) {
            for (uint64_t i = in.up_off[r]; i < in.up_off[r + 1]; ++i)
                if (in.up_w[i] != INF && !is_anc(in.up_tgt[i], r)) flag("up arc between non-ancestors");
            for (uint64_t i = in.dn_off[r]; i < in.dn_off[r + 1]; ++i)
                if (in.dn_w[i] != INF && !is_anc(r, in.dn_tgt[i])) flag("down arc between non-ancestors");
        }
    });
    if (failed()) return false;
    // stage of a position: segment level for top nodes, then the bottom levels
    std::vector<uint32_t> seg_of_pos(L.n_top), lvl_of_seg(L.seg_pos.size() - 1), stage(n);
    for (size_t l = 0; l + 1 < L.seglvl_seg.size(); ++l)
        for (uint32_t s = L.seglvl_seg[l]; s < L.seglvl_seg[l + 1]; ++s) lvl_of_seg[s] = (uint32_t)l;
    for (size_t s = 0; s + 1 < L.seg_pos.size(); ++s)
        for (uint32_t p = L.seg_pos[s]; p < L.seg_pos[s + 1]; ++p) {
            seg_of_pos[p] = (uint32_t)s;
            stage[p] = lvl_of_seg[s];
        }
    const uint32_t nsl = (uint32_t)L.seglvl_seg.size() - 1;
    for (size_t b = 0; b + 1 < L.botlvl_pos.size(); ++b)
        for (uint32_t p = L.botlvl_pos[b]; p < L.botlvl_pos[b + 1]; ++p) stage[p] = nsl + (uint32_t)b;
    parallel_blocks(n, 4096, [&](size_t pb, size_t pe) {
        for (uint32_t p = (uint32_t)pb; p < pe; ++p) {
            const bool tp = p < L.n_top;
            if (L.in_split[p] < L.in_off[p] || L.in_split[p] > L.in_off[p + 1]) { flag("bad in_split"); return; }
            for (uint32_t k = L.in_off[p]; k < L.in_off[p + 1]; ++k) {
                const uint32_t q = L.in_src[k];
                const bool second = k >= L.in_split[p];
                bool ok;
                if (tp) ok = second ? (q < L.n_top && seg_of_pos[q] == seg_of_pos[p] && q < p)
                                    : (q < n && stage[q] < stage[p]);
                else ok = second ? q < L.n_top : (q >= L.n_top && q < n && stage[q] < stage[p]);
                if (!ok) flag("down schedule violates an arc dependency");
            }
            if (L.up_split[p] < L.up_off[p] || L.up_split[p] > L.up_off[p + 1]) { flag("bad up_split"); return; }
            for (uint32_t k = L.up_off[p]; k < L.up_off[p + 1]; ++k) {
                if (k >= L.up_split[p]) {  // intra-segment
                    if (!tp) { flag("intra up arc at a bottom node"); return; }
                    const uint32_t q = L.seg_pos[seg_of_pos[p]] + L.up_tgt[k];
                    if (!(q < p && seg_of_pos[q] == seg_of_pos[p])) flag("bad intra up arc");
                } else {
                    const uint32_t q = L.up_tgt[k];
                    if (q >= n) { flag("up target out of range"); return; }
                    if (tp && !(q < L.n_top && stage[q] < stage[p])) flag("up schedule violates an arc dependency");
                    if (!tp && !(stage[q] < stage[p])) flag("bottom up arc to a non-ancestor level");
                }
            }
            if (L.pos_parent[p] != NONE && L.pos_parent[p] >= n) flag("bad parent position");
        }
    });
    if (failed()) return false;
    for (size_t s = 0; s + 1 < L.seg_pos.size(); ++s)
        if (L.seg_pos[s + 1] <= L.seg_pos[s]) { err = "empty segment"; return false; }
    if (L.in_dd.size() != L.in_off[n] - L.in_off[L.n_top]) { err = "depth gap table disagrees"; return false; }
    parallel_blocks(n - L.n_top, 4096, [&](size_t b, size_t e) {
        for (uint32_t p = L.n_top + (uint32_t)b; p < L.n_top + e; ++p)
            for (uint32_t a = L.in_off[p]; a < L.in_split[p]; ++a) {
                const uint32_t u = L.in_src[a], dd = L.in_dd[a - L.in_off[L.n_top]];
                // the source is the ancestor dd levels up: follow the parents
                uint32_t x = p;
                for (uint32_t k = 0; k < dd && x != NONE; ++k) x = L.pos_parent[x];
                if (dd == 0 || x != u) flag("depth gap of a bottom arc");
            }
    });
    if (failed()) return false;
    // bottom chunks: every bottom position once, sources in the top region
    {
        std::vector<uint8_t> seen(n, 0);
        const size_t nch = L.ch_off.size() - 1;
        if (L.cs_off.size() != nch + 1) { err = "chunk tables disagree"; return false; }
        for (size_t c = 0; c < nch; ++c) {
            const uint32_t len = L.ch_off[c + 1] - L.ch_off[c];
            if (len == 0 || len > 0xFFFF) { err = "bad chunk size"; return false; }
            for (uint32_t i = L.ch_off[c]; i < L.ch_off[c + 1]; ++i) {
                const uint32_t p = L.ch_pos[i];
                if (p < L.n_top || p >= n || seen[p]) { err = "chunk positions not a bottom partition"; return false; }
                seen[p] = 1;
            }
            for (uint32_t k = L.cs_off[c]; k < L.cs_off[c + 1]; ++k) {
                if (L.cs_src[k] >= L.n_top) { err = "chunk source outside the top region"; return false; }
                for (uint32_t a = L.cs_aoff[k]; a < L.cs_aoff[k + 1]; ++a)
                    if (L.ca_v[a] >= len) { err = "chunk arc head out of chunk"; return false; }
            }
        }
        if (L.ch_pos.size() != n - L.n_top) { err = "chunks miss bottom positions"; return false; }
    }
    // dense tiles: the finite entries of each table are exactly the arcs they
    // replace (parallel arcs merged to their minimum), keys are readable and
    // final before the segment's stage
    {
        using T3 = std::tuple<uint32_t, uint32_t, uint32_t>;
        auto merged = [](std::vector<T3> &v) {  // sort, keep the minimum weight per (a, b)
            std::sort(v.begin(), v.end());
            size_t o = 0;
            for (size_t i = 0; i < v.size(); ++i)
                if (o == 0 || std::get<0>(v[i]) != std::get<0>(v[o - 1]) || std::get<1>(v[i]) != std::get<1>(v[o - 1]))
                    v[o++] = v[i];
            v.resize(o);
        };
        const uint32_t nseg = (uint32_t)L.seg_pos.size() - 1;
        const size_t nblk = L.pt_blk.empty() ? 0 : L.pt_blk.back();
        if (L.xr_off.size() != (size_t)nseg + 1 || L.pt_blk.size() != nseg + 1 ||
            L.tb_dn.size() != 64 * (size_t)nseg || L.tb_up.size() != 64 * (size_t)nseg ||
            L.xr_run.size() != 8 * (size_t)nseg || L.xw_off.size() != (size_t)nseg + 1 ||
            L.xr_off[nseg] != L.xr_src.size() ||
            L.xw_off[nseg] != L.xr_w.size() || L.pr_off.size() != nblk + 1 || L.pw_off.size() != nblk + 1 ||
            L.pr_wid.size() != nblk || L.pr_off.back() != L.pr_x.size() || L.pw_off.back() != L.pr_w.size() ||
            L.pt_tgt.size() != 64 * nblk) {
            err = "dense tile tables disagree";
            return false;
        }
        for (uint32_t s = 0; s < nseg; ++s) {
            const uint32_t *run = &L.xr_run[8 * (size_t)s];
            uint64_t weights = 0;
            for (uint32_t c = 0, st = 0; c < 8; st = run[c++]) {
                if (run[c] < st) { err = "row runs of a segment out of order"; return false; }
                weights += (uint64_t)(run[c] - st) * 8 * (8 - c);
            }
            if (run[7] != L.xr_off[s + 1] - L.xr_off[s] || (uint64_t)L.xw_off[s + 1] - L.xw_off[s] != weights) {
                err = "row runs disagree with a segment's rows";
                return false;
            }
        }
        for (size_t k = 0; k < nblk; ++k) {
            const uint32_t W = L.pr_wid[k];
            if ((W != 16 && W != 32 && W != 64) ||
                (uint64_t)L.pw_off[k + 1] - L.pw_off[k] != (uint64_t)W * (L.pr_off[k + 1] - L.pr_off[k])) {
                err = "bad row width of a target block";
                return false;
            }
            for (uint32_t j = W; j < 64; ++j)
                if (L.pt_tgt[64 * k + j] != NONE) { err = "target beyond a block's row width"; return false; }
        }
        parallel_blocks(nseg, 16, [&](size_t sb, size_t se) {
            std::vector<T3> want, got;
            for (uint32_t s = (uint32_t)sb; s < se; ++s) {
                const uint32_t p0 = L.seg_pos[s], len = L.seg_pos[s + 1] - p0;
                if (len > 64) { flag("segment longer than a tile"); return; }
                // DOWN from outside: (head offset, source, w)
                want.clear();
                got.clear();
                for (uint32_t i = 0; i < len; ++i)
                    for (uint32_t k = L.in_off[p0 + i]; k < L.in_split[p0 + i]; ++k)
                        want.emplace_back(i, L.in_src[k], L.in_w[k]);
                size_t wo = L.xw_off[s];
                for (uint32_t r = L.xr_off[s]; r < L.xr_off[s + 1]; ++r) {
                    const uint32_t u = L.xr_src[r], i = r - L.xr_off[s];
                    uint32_t c = 0;
                    while (i >= L.xr_run[8 * (size_t)s + c]) ++c;  // i < run[7]: ends
                    const uint32_t W = 8 * (8 - c);
                    if (u >= n || stage[u] >= stage[p0]) { flag("tile source not final"); return; }
                    for (uint32_t j = 0; j < W; ++j) {
                        const uint32_t w = L.xr_w[wo + j];
                        if (w == INF) continue;
                        if (j >= len) { flag("tile head beyond the segment"); return; }
                        got.emplace_back(j, u, w);
                    }
                    wo += W;
                }
                merged(want);
                std::sort(got.begin(), got.end());
                if (want != got) { flag("DOWN tiles differ from the arcs they replace"); return; }
                // intra: (tail, head, w) for both directions
                for (int dir = 0; dir < 2; ++dir) {
                    want.clear();
                    got.clear();
                    for (uint32_t i = 0; i < len; ++i) {
                        if (dir == 0)
                            for (uint32_t k = L.in_split[p0 + i]; k < L.in_off[p0 + i + 1]; ++k)
                                want.emplace_back(L.in_src[k] - p0, i, L.in_w[k]);
                        else
                            for (uint32_t k = L.up_split[p0 + i]; k < L.up_off[p0 + i + 1]; ++k)
                                want.emplace_back(i, L.up_tgt[k], L.up_w[k]);
                    }
                    const auto &tab = dir == 0 ? L.tb_dn : L.tb_up;
                    for (uint32_t a = 0; a < 8; ++a)
                        for (uint32_t o = 0; o < 8; ++o) {
                            const uint32_t k = tab[64 * (size_t)s + 8 * a + o];
                            if (k == NONE) continue;
                            if ((dir == 0 && a > o) || (dir == 1 && a < o) || 64 * (size_t)k + 64 > L.tb_w.size()) {
                                flag("bad intra block");
                                return;
                            }
                            for (uint32_t i = 0; i < 8; ++i)
                                for (uint32_t j = 0; j < 8; ++j) {
                                    const uint32_t w = L.tb_w[64 * (size_t)k + 8 * i + j];
                                    if (w == INF) continue;
                                    const uint32_t t = 8 * a + i, h = 8 * o + j;
                                    if (t >= len || h >= len || (dir == 0 ? t >= h : t <= h)) {
                                        flag("intra block arc against the sweep order");
                                        return;
                                    }
                                    got.emplace_back(t, h, w);
                                }
                        }
                    merged(want);
                    std::sort(got.begin(), got.end());
                    if (want != got) { flag("intra blocks differ from the arcs they replace"); return; }
                }
                // UP leaving: (target, tail position, w)
                want.clear();
                got.clear();
                for (uint32_t i = 0; i < len; ++i)
                    for (uint32_t k = L.up_off[p0 + i]; k < L.up_split[p0 + i]; ++k)
                        want.emplace_back(L.up_tgt[k], p0 + i, L.up_w[k]);
                for (uint32_t kb = L.pt_blk[s]; kb < L.pt_blk[s + 1]; ++kb)
                    for (uint32_t r = L.pr_off[kb]; r < L.pr_off[kb + 1]; ++r) {
                        const uint32_t x = L.pr_x[r], W = L.pr_wid[kb];
                        if (x < p0 || x >= p0 + len) { flag("push tail outside the segment"); return; }
                        for (uint32_t j = 0; j < W; ++j) {
                            const uint32_t w = L.pr_w[L.pw_off[kb] + (size_t)(r - L.pr_off[kb]) * W + j];
                            const uint32_t t = L.pt_tgt[64 * (size_t)kb + j];
                            if (w == INF) continue;
                            if (t == NONE || t >= L.n_top || stage[t] >= stage[p0]) {
                                flag("push target not an earlier top node");
                                return;
                            }
                            got.emplace_back(t, x, w);
                        }
                    }
                merged(want);
                std::sort(got.begin(), got.end());
                if (want != got) { flag("UP push tiles differ from the arcs they replace"); return; }
            }
        });
        if (failed()) return false;
        if (L.pk_seg.size() != L.pt_blk[nseg]) { err = "target block table disagrees"; return false; }
        for (uint32_t s = 0; s < nseg; ++s)
            for (uint32_t k = L.pt_blk[s]; k < L.pt_blk[s + 1]; ++k)
                if (L.pk_seg[k] != s) { err = "target block of another segment"; return false; }
    }
    return true;
}

}  // namespace bfg
