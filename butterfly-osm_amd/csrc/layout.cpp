// layout.cpp — see layout.hpp.
#include "layout.hpp"

#include <algorithm>
#include <numeric>
#include <tuple>

namespace bfg {

namespace {
template <class T>
void prefix(std::vector<T> &v);

struct Arc3 {
    uint32_t key, col, w;  // key: source (or tail), col: column in a 64-wide block
};

// Appends the rows of one 64-column block from arcs (key, column, weight):
// one row per distinct key, ascending, holding the key and the weights of its
// arcs to the 64 columns (INF where there is none; parallel arcs keep the
// minimum).  Returns the number of rows.
uint32_t build_rows(std::vector<Arc3> &arcs, std::vector<uint32_t> &rkey, std::vector<uint32_t> &rw) {
    std::sort(arcs.begin(), arcs.end(), [](const Arc3 &a, const Arc3 &b) {
        return a.key != b.key ? a.key < b.key : a.col < b.col;
    });
    uint32_t rows = 0;
    for (size_t i = 0; i < arcs.size();) {
        const uint32_t key = arcs[i].key;
        const size_t base = rw.size();
        rw.resize(base + 64, INF);
        for (; i < arcs.size() && arcs[i].key == key; ++i)
            rw[base + arcs[i].col] = std::min(rw[base + arcs[i].col], arcs[i].w);
        rkey.push_back(key);
        ++rows;
    }
    return rows;
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
    uint64_t mu = 0, md = 0;
    for (uint32_t r = 0; r < n; ++r) {
        for (uint64_t i = in.up_off[r]; i < in.up_off[r + 1]; ++i) {
            uint32_t v = in.up_tgt[i];
            if (v >= n || v <= r) { err = "UP arc target not above its tail (not a rank-aligned CCH)"; return false; }
            if (in.up_w[i] != INF) ++mu; else ++L.m_up_inf;
        }
        for (uint64_t i = in.dn_off[r]; i < in.dn_off[r + 1]; ++i) {
            uint32_t v = in.dn_tgt[i];
            if (v >= r) { err = "DOWN arc target not below its tail (not a rank-aligned CCH)"; return false; }
            if (in.dn_w[i] != INF) ++md; else ++L.m_down_inf;
        }
    }
    if (mu >= 0xFFFFFFFFull || md >= 0xFFFFFFFFull) { err = "more than 2^32-1 finite arcs"; return false; }
    L.m_up = mu;
    L.m_down = md;
    // UP arcs by tail rank, duplicates (tail, target) merged to the minimum weight
    std::vector<uint32_t> uo(n + 1, 0), ut, uw;
    ut.reserve(mu);
    uw.reserve(mu);
    {
        std::vector<std::pair<uint32_t, uint32_t>> tmp;
        for (uint32_t r = 0; r < n; ++r) {
            uo[r] = (uint32_t)ut.size();
            tmp.clear();
            for (uint64_t i = in.up_off[r]; i < in.up_off[r + 1]; ++i)
                if (in.up_w[i] != INF) tmp.emplace_back(in.up_tgt[i], in.up_w[i]);
            std::sort(tmp.begin(), tmp.end());
            for (size_t i = 0; i < tmp.size(); ++i) {
                if (i && tmp[i].first == tmp[i - 1].first) continue;
                ut.push_back(tmp[i].first);
                uw.push_back(tmp[i].second);
            }
        }
        uo[n] = (uint32_t)ut.size();
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
    std::vector<uint32_t> seg_of(n, NONE);
    for (uint32_t v = 0; v < n; ++v)
        if (top[v]) seg_of[v] = chain_seg0[chain_of[v]] + idx_in_chain[v] / SEG;
    // bottom heights
    std::vector<uint32_t> height(n, 0);
    for (uint32_t v = 0; v < n; ++v) {
        uint32_t p = L.parent[v];
        if (!top[v] && p != NONE && !top[p]) height[p] = std::max(height[p], height[v] + 1);
    }
    for (uint32_t v = 0; v < n; ++v)
        if (!top[v]) L.max_height = std::max(L.max_height, height[v]);
    // ---- positions ----
    uint32_t n_seglvl = 0;
    for (uint32_t l : seg_level) n_seglvl = std::max(n_seglvl, l + 1);
    std::vector<std::vector<uint32_t>> lvl_segs(n_seglvl);
    for (uint32_t s = 0; s < seg_level.size(); ++s) lvl_segs[seg_level[s]].push_back(s);
    L.pos_of_rank.assign(n, NONE);
    L.rank_of_pos.assign(n, NONE);
    uint32_t pos = 0;
    std::vector<uint32_t> seg_start(seg_level.size());
    for (uint32_t l = 0; l < n_seglvl; ++l) {
        L.seglvl_seg.push_back((uint32_t)L.seg_pos.size());
        for (uint32_t s : lvl_segs[l]) {
            L.seg_pos.push_back(pos);
            seg_start[s] = pos;
            const auto &ch = chains[seg_range[s].first];
            uint32_t i0 = seg_range[s].second, i1 = std::min<uint32_t>((uint32_t)ch.size(), i0 + SEG);
            L.max_segment = std::max(L.max_segment, i1 - i0);
            for (uint32_t i = i0; i < i1; ++i) { L.pos_of_rank[ch[i]] = pos; L.rank_of_pos[pos] = ch[i]; ++pos; }
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
        for (uint32_t r = 0; r < n; ++r)
            for (uint64_t i = in.dn_off[r]; i < in.dn_off[r + 1]; ++i)
                if (in.dn_w[i] != INF) cnt[L.pos_of_rank[in.dn_tgt[i]]]++;
        L.in_off = cnt;
        prefix(L.in_off);
        L.in_src.resize(md);
        L.in_w.resize(md);
        std::vector<uint32_t> cur(L.in_off.begin(), L.in_off.end() - 1);
        for (uint32_t r = 0; r < n; ++r) {
            uint32_t pr = L.pos_of_rank[r];
            for (uint64_t i = in.dn_off[r]; i < in.dn_off[r + 1]; ++i) {
                if (in.dn_w[i] == INF) continue;
                uint32_t k = cur[L.pos_of_rank[in.dn_tgt[i]]]++;
                L.in_src[k] = pr;
                L.in_w[k] = in.dn_w[i];
            }
        }
        L.in_split.resize(n);
        std::vector<std::pair<uint32_t, uint32_t>> tmp;
        for (uint32_t p = 0; p < n; ++p) {
            uint32_t a = L.in_off[p], b = L.in_off[p + 1];
            uint32_t v = L.rank_of_pos[p];
            // second class: top head — arcs from its own segment; bottom head —
            // arcs from top sources
            auto second = [&](uint32_t src_pos) {
                return top[v] ? same_seg(L.rank_of_pos[src_pos], v) : src_pos < L.n_top;
            };
            tmp.resize(b - a);
            for (uint32_t k = a; k < b; ++k) tmp[k - a] = {L.in_src[k], L.in_w[k]};
            std::sort(tmp.begin(), tmp.end(), [&](const std::pair<uint32_t, uint32_t> &x,
                                                  const std::pair<uint32_t, uint32_t> &y) {
                bool ix = second(x.first), iy = second(y.first);
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
            if (top[v]) L.m_down_intra += b - split;
        }
    }
    // ---- UP by tail position, outside arcs first ----
    {
        L.up_off.assign(n + 1, 0);
        for (uint32_t r = 0; r < n; ++r) L.up_off[L.pos_of_rank[r]] = uo[r + 1] - uo[r];
        prefix(L.up_off);
        L.up_tgt.resize(ut.size());
        L.up_w.resize(ut.size());
        L.up_split.resize(n);
        std::vector<std::pair<uint32_t, uint32_t>> outs, ins;
        for (uint32_t p = 0; p < n; ++p) {
            uint32_t x = L.rank_of_pos[p];
            outs.clear();
            ins.clear();
            for (uint32_t k = uo[x]; k < uo[x + 1]; ++k) {
                uint32_t y = ut[k];
                if (same_seg(x, y)) ins.emplace_back(L.pos_of_rank[y] - seg_start[seg_of[y]], uw[k]);
                else outs.emplace_back(L.pos_of_rank[y], uw[k]);
            }
            std::sort(outs.begin(), outs.end());
            std::sort(ins.begin(), ins.end());
            uint32_t k = L.up_off[p];
            for (auto &e : outs) { L.up_tgt[k] = e.first; L.up_w[k] = e.second; ++k; }
            L.up_split[p] = k;
            for (auto &e : ins) { L.up_tgt[k] = e.first; L.up_w[k] = e.second; ++k; }
            L.m_up_intra += ins.size();
        }
    }
    // ---- grouped DOWN arcs: segments (outside arcs) and bottom chunks (top-sourced arcs) ----
    const uint32_t BB = std::max<uint32_t>(lp.block_limit, 1);
    // groups arcs (source, local head, weight) of one node set by source
    auto group = [](std::vector<std::tuple<uint32_t, uint32_t, uint32_t>> &arcs, std::vector<uint32_t> &src,
                    std::vector<uint32_t> &aoff, std::vector<uint16_t> &av, std::vector<uint32_t> &aw) {
        std::sort(arcs.begin(), arcs.end());
        for (size_t i = 0; i < arcs.size(); ++i) {
            if (i == 0 || std::get<0>(arcs[i]) != std::get<0>(arcs[i - 1])) {
                src.push_back(std::get<0>(arcs[i]));
                aoff.push_back((uint32_t)av.size());
            }
            av.push_back((uint16_t)std::get<1>(arcs[i]));
            aw.push_back(std::get<2>(arcs[i]));
        }
    };
    std::vector<std::tuple<uint32_t, uint32_t, uint32_t>> arcs;
    {
        const uint32_t nseg = (uint32_t)L.seg_pos.size() - 1;
        L.bs_off.assign(nseg + 1, 0);
        for (uint32_t s = 0; s < nseg; ++s) {
            L.bs_off[s] = (uint32_t)L.bs_src.size();
            const uint32_t p0 = L.seg_pos[s], p1 = L.seg_pos[s + 1];
            L.max_block = std::max(L.max_block, p1 - p0);
            arcs.clear();
            for (uint32_t p = p0; p < p1; ++p)
                for (uint32_t k = L.in_off[p]; k < L.in_split[p]; ++k) arcs.emplace_back(L.in_src[k], p - p0, L.in_w[k]);
            group(arcs, L.bs_src, L.bs_aoff, L.ba_v, L.ba_w);
        }
        L.bs_off[nseg] = (uint32_t)L.bs_src.size();
        L.bs_aoff.push_back((uint32_t)L.ba_v.size());
    }
    {
        std::vector<uint32_t> bot;
        for (uint32_t p = L.n_top; p < n; ++p) bot.push_back(p);
        std::sort(bot.begin(), bot.end(), [&](uint32_t a, uint32_t b) { return L.pos_pre[a] < L.pos_pre[b]; });
        L.ch_off.clear();
        L.ch_pos.clear();
        for (size_t i = 0; i < bot.size(); i += BB) {
            L.ch_off.push_back((uint32_t)L.ch_pos.size());
            const size_t e = std::min(bot.size(), i + BB);
            arcs.clear();
            L.cs_off.push_back((uint32_t)L.cs_src.size());
            for (size_t k = i; k < e; ++k) {
                const uint32_t p = bot[k];
                L.ch_pos.push_back(p);
                for (uint32_t a = L.in_split[p]; a < L.in_off[p + 1]; ++a)
                    arcs.emplace_back(L.in_src[a], (uint32_t)(k - i), L.in_w[a]);
            }
            group(arcs, L.cs_src, L.cs_aoff, L.ca_v, L.ca_w);
        }
        L.ch_off.push_back((uint32_t)L.ch_pos.size());
        L.cs_off.push_back((uint32_t)L.cs_src.size());
        L.cs_aoff.push_back((uint32_t)L.ca_v.size());
    }
    // UP arcs leaving each segment, grouped by target position
    {
        const uint32_t nseg = (uint32_t)L.seg_pos.size() - 1;
        L.ut_off.assign(nseg + 1, 0);
        L.ut_tgt.clear();
        L.ut_aoff.clear();
        L.uta_x.clear();
        L.uta_w.clear();
        std::vector<std::tuple<uint32_t, uint32_t, uint32_t>> arcs;  // (target, local tail, weight)
        for (uint32_t s = 0; s < nseg; ++s) {
            L.ut_off[s] = (uint32_t)L.ut_tgt.size();
            const uint32_t p0 = L.seg_pos[s], p1 = L.seg_pos[s + 1];
            arcs.clear();
            for (uint32_t p = p0; p < p1; ++p)
                for (uint32_t k = L.up_off[p]; k < L.up_split[p]; ++k) arcs.emplace_back(L.up_tgt[k], p - p0, L.up_w[k]);
            std::sort(arcs.begin(), arcs.end());
            for (size_t i = 0; i < arcs.size(); ++i) {
                if (i == 0 || std::get<0>(arcs[i]) != std::get<0>(arcs[i - 1])) {
                    L.ut_tgt.push_back(std::get<0>(arcs[i]));
                    L.ut_aoff.push_back((uint32_t)L.uta_x.size());
                }
                L.uta_x.push_back((uint16_t)std::get<1>(arcs[i]));
                L.uta_w.push_back(std::get<2>(arcs[i]));
            }
        }
        L.ut_off[nseg] = (uint32_t)L.ut_tgt.size();
        L.ut_aoff.push_back((uint32_t)L.uta_x.size());
    }
    // intra-segment arc lists by tail
    {
        const uint32_t nt = L.n_top;
        L.sd_off.assign(nt + 1, 0);
        L.su_off.assign(nt + 1, 0);
        for (uint32_t p = 0; p < nt; ++p) {
            for (uint32_t k = L.in_split[p]; k < L.in_off[p + 1]; ++k) L.sd_off[L.in_src[k]]++;
            L.su_off[p] = L.up_off[p + 1] - L.up_split[p];
        }
        prefix(L.sd_off);
        prefix(L.su_off);
        L.sd_tgt.resize(L.sd_off[nt]);
        L.sd_w.resize(L.sd_off[nt]);
        L.su_tgt.resize(L.su_off[nt]);
        L.su_w.resize(L.su_off[nt]);
        std::vector<uint32_t> cur(L.sd_off.begin(), L.sd_off.end() - 1);
        for (size_t s = 0; s + 1 < L.seg_pos.size(); ++s)
            for (uint32_t p = L.seg_pos[s]; p < L.seg_pos[s + 1]; ++p) {  // heads ascending per tail
                for (uint32_t k = L.in_split[p]; k < L.in_off[p + 1]; ++k) {
                    const uint32_t d = cur[L.in_src[k]]++;
                    L.sd_tgt[d] = (uint16_t)(p - L.seg_pos[s]);
                    L.sd_w[d] = L.in_w[k];
                }
                for (uint32_t k = L.up_split[p], d = L.su_off[p]; k < L.up_off[p + 1]; ++k, ++d) {
                    L.su_tgt[d] = (uint16_t)L.up_tgt[k];
                    L.su_w[d] = L.up_w[k];
                }
            }
    }
    // ---- dense tiles of the top region (segments of at most 64 nodes) ----
    {
        const uint32_t nseg = (uint32_t)L.seg_pos.size() - 1;
        if (L.max_segment > 64) { err = "segment_limit above 64"; return false; }
        L.xr_off.assign(nseg + 1, 0);
        L.tb_dn.assign(64 * (size_t)nseg, NONE);
        L.tb_up.assign(64 * (size_t)nseg, NONE);
        L.pt_blk.assign(nseg + 1, 0);
        L.xr_src.clear(); L.xr_w.clear(); L.tb_w.clear();
        L.pt_tgt.clear(); L.pr_off.clear(); L.pr_x.clear(); L.pr_w.clear();
        std::vector<Arc3> a3;
        std::vector<std::tuple<uint32_t, uint32_t, uint32_t>> outs;  // (target, tail position, weight)
        std::vector<uint32_t> colv;                                   // target column of each arc of outs
        uint64_t xrows = 0, prows = 0;
        auto put = [&](std::vector<uint32_t> &tab, uint32_t s, uint32_t a, uint32_t o, uint32_t i, uint32_t j,
                       uint32_t w) {
            uint32_t &k = tab[64 * (size_t)s + 8 * a + o];
            if (k == NONE) {
                k = (uint32_t)(L.tb_w.size() / 64);
                L.tb_w.resize(L.tb_w.size() + 64, INF);
            }
            uint32_t &slot = L.tb_w[64 * (size_t)k + 8 * i + j];
            slot = std::min(slot, w);
        };
        for (uint32_t s = 0; s < nseg; ++s) {
            const uint32_t p0 = L.seg_pos[s], len = L.seg_pos[s + 1] - p0;
            L.xr_off[s] = (uint32_t)xrows;
            a3.clear();
            for (uint32_t j = 0; j < len; ++j)
                for (uint32_t k = L.in_off[p0 + j]; k < L.in_split[p0 + j]; ++k) a3.push_back({L.in_src[k], j, L.in_w[k]});
            xrows += build_rows(a3, L.xr_src, L.xr_w);
            for (uint32_t i = 0; i < len; ++i) {
                const uint32_t p = p0 + i;
                for (uint32_t k = L.in_split[p]; k < L.in_off[p + 1]; ++k) {
                    const uint32_t ti = L.in_src[k] - p0;  // tail before head: ti < i
                    put(L.tb_dn, s, ti / 8, i / 8, ti % 8, i % 8, L.in_w[k]);
                }
                for (uint32_t k = L.up_split[p]; k < L.up_off[p + 1]; ++k) {
                    const uint32_t t = L.up_tgt[k];  // head offset, t < i
                    put(L.tb_up, s, i / 8, t / 8, i % 8, t % 8, L.up_w[k]);
                }
            }
            outs.clear();
            for (uint32_t i = 0; i < len; ++i)
                for (uint32_t k = L.up_off[p0 + i]; k < L.up_split[p0 + i]; ++k)
                    outs.emplace_back(L.up_tgt[k], p0 + i, L.up_w[k]);
            std::sort(outs.begin(), outs.end());
            L.pt_blk[s] = (uint32_t)(L.pt_tgt.size() / 64);
            const size_t t0 = L.pt_tgt.size();
            for (size_t q = 0; q < outs.size(); ++q)
                if (q == 0 || std::get<0>(outs[q]) != std::get<0>(outs[q - 1])) L.pt_tgt.push_back(std::get<0>(outs[q]));
            const uint32_t nt = (uint32_t)(L.pt_tgt.size() - t0);
            L.pt_tgt.resize(t0 + (size_t)(nt + 63) / 64 * 64, NONE);
            colv.resize(outs.size());
            for (size_t q = 0, c = 0; q < outs.size(); ++q) {
                if (q > 0 && std::get<0>(outs[q]) != std::get<0>(outs[q - 1])) ++c;
                colv[q] = (uint32_t)c;
            }
            size_t q = 0;
            for (uint32_t kb = 0; kb * 64 < nt; ++kb) {
                L.pr_off.push_back((uint32_t)prows);
                a3.clear();
                for (; q < outs.size() && colv[q] < 64 * (kb + 1); ++q)
                    a3.push_back({std::get<1>(outs[q]), colv[q] - 64 * kb, std::get<2>(outs[q])});
                prows += build_rows(a3, L.pr_x, L.pr_w);
            }
        }
        L.xr_off[nseg] = (uint32_t)xrows;
        L.pt_blk[nseg] = (uint32_t)(L.pt_tgt.size() / 64);
        L.pr_off.push_back((uint32_t)prows);
        const uint32_t R = std::max<uint32_t>(lp.ell_chunk, 1);
        L.du_lvl.clear(); L.du_seg.clear(); L.du_r0.clear(); L.du_r1.clear();
        for (size_t l = 0; l + 1 < L.seglvl_seg.size(); ++l) {
            L.du_lvl.push_back((uint32_t)L.du_seg.size());
            for (uint32_t s = L.seglvl_seg[l]; s < L.seglvl_seg[l + 1]; ++s)
                for (uint32_t r = L.xr_off[s]; r < L.xr_off[s + 1]; r += R) {
                    L.du_seg.push_back(s);
                    L.du_r0.push_back(r);
                    L.du_r1.push_back(std::min(r + R, L.xr_off[s + 1]));
                }
        }
        L.du_lvl.push_back((uint32_t)L.du_seg.size());
        L.pk_seg.assign(L.pt_blk[nseg], 0);
        for (uint32_t s = 0; s < nseg; ++s)
            for (uint32_t k = L.pt_blk[s]; k < L.pt_blk[s + 1]; ++k) L.pk_seg[k] = s;
        L.ell_down_entries = L.xr_w.size();
        L.ell_up_entries = L.pr_w.size();
        if (L.xr_w.size() >= (1ull << 32) || L.pr_w.size() >= (1ull << 32) || L.tb_w.size() >= (1ull << 32)) {
            err = "dense tiles exceed 2^32 weights";
            return false;
        }
    }
    return true;
}

bool check_layout(const CchInput &in, const Layout &L, std::string &err) {
    const uint32_t n = in.n;
    auto is_anc = [&](uint32_t a, uint32_t d) {  // rank a ancestor-or-self of rank d
        return L.pre[a] <= L.pre[d] && L.pre[d] < L.pre[a] + L.subtree[a];
    };
    for (uint32_t r = 0; r < n; ++r) {
        for (uint64_t i = in.up_off[r]; i < in.up_off[r + 1]; ++i)
            if (in.up_w[i] != INF && !is_anc(in.up_tgt[i], r)) { err = "up arc between non-ancestors"; return false; }
        for (uint64_t i = in.dn_off[r]; i < in.dn_off[r + 1]; ++i)
            if (in.dn_w[i] != INF && !is_anc(r, in.dn_tgt[i])) { err = "down arc between non-ancestors"; return false; }
    }
    // stage of a position: segment level for top nodes, then the bottom levels
    std::vector<uint32_t> seg_of_pos(L.n_top), lvl_of_seg(L.seg_pos.size() - 1), stage(n);
    for (size_t l = 0; l + 1 < L.seglvl_seg.size(); ++l)
        for (uint32_t s = L.seglvl_seg[l]; s < L.seglvl_seg[l + 1]; ++s) lvl_of_seg[s] = (uint32_t)l;
    for (size_t s = 0; s + 1 < L.seg_pos.size(); ++s)
        for (uint32_t p = L.seg_pos[s]; p < L.seg_pos[s + 1]; ++p) { seg_of_pos[p] = (uint32_t)s; stage[p] = lvl_of_seg[s]; }
    const uint32_t nsl = (uint32_t)L.seglvl_seg.size() - 1;
    for (size_t b = 0; b + 1 < L.botlvl_pos.size(); ++b)
        for (uint32_t p = L.botlvl_pos[b]; p < L.botlvl_pos[b + 1]; ++p) stage[p] = nsl + (uint32_t)b;
    for (uint32_t p = 0; p < n; ++p) {
        const bool tp = p < L.n_top;
        if (L.in_split[p] < L.in_off[p] || L.in_split[p] > L.in_off[p + 1]) { err = "bad in_split"; return false; }
        for (uint32_t k = L.in_off[p]; k < L.in_off[p + 1]; ++k) {
            uint32_t q = L.in_src[k];
            bool second = k >= L.in_split[p];
            bool ok;
            if (tp) ok = second ? (q < L.n_top && seg_of_pos[q] == seg_of_pos[p] && q < p)
                                : (q < n && stage[q] < stage[p]);
            else ok = second ? q < L.n_top : (q >= L.n_top && q < n && stage[q] < stage[p]);
            if (!ok) { err = "down schedule violates an arc dependency"; return false; }
        }
        if (L.up_split[p] < L.up_off[p] || L.up_split[p] > L.up_off[p + 1]) { err = "bad up_split"; return false; }
        for (uint32_t k = L.up_off[p]; k < L.up_off[p + 1]; ++k) {
            bool intra = k >= L.up_split[p];
            if (intra) {
                if (!tp) { err = "intra up arc at a bottom node"; return false; }
                uint32_t q = L.seg_pos[seg_of_pos[p]] + L.up_tgt[k];
                if (!(q < p && seg_of_pos[q] == seg_of_pos[p])) { err = "bad intra up arc"; return false; }
            } else {
                uint32_t q = L.up_tgt[k];
                if (q >= n) { err = "up target out of range"; return false; }
                if (tp && !(q < L.n_top && stage[q] < stage[p])) { err = "up schedule violates an arc dependency"; return false; }
                if (!tp && !(stage[q] < stage[p])) { err = "bottom up arc to a non-ancestor level"; return false; }
            }
        }
        if (L.pos_parent[p] != NONE && L.pos_parent[p] >= n) { err = "bad parent position"; return false; }
    }
    // grouped segment arcs: sources final before the segment's stage
    for (size_t s = 0; s + 1 < L.seg_pos.size(); ++s) {
        const uint32_t p0 = L.seg_pos[s], p1 = L.seg_pos[s + 1];
        if (p1 <= p0 || p1 - p0 > 0xFFFF) { err = "bad segment size"; return false; }
        for (uint32_t p = p0; p < p1; ++p) {
            for (uint32_t k = L.sd_off[p]; k < L.sd_off[p + 1]; ++k)
                if (!(L.sd_tgt[k] > p - p0 && L.sd_tgt[k] < p1 - p0)) { err = "intra DOWN arc not to a later node"; return false; }
            for (uint32_t k = L.su_off[p]; k < L.su_off[p + 1]; ++k)
                if (!(L.su_tgt[k] < p - p0)) { err = "intra UP arc not to an earlier node"; return false; }
        }
        for (uint32_t k = L.bs_off[s]; k < L.bs_off[s + 1]; ++k) {
            if (!(L.bs_src[k] < n && stage[L.bs_src[k]] < stage[p0])) { err = "segment source not final"; return false; }
            for (uint32_t a = L.bs_aoff[k]; a < L.bs_aoff[k + 1]; ++a)
                if (L.ba_v[a] >= p1 - p0) { err = "segment arc head out of segment"; return false; }
        }
    }
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
    for (size_t s = 0; s + 1 < L.seg_pos.size(); ++s) {
        const uint32_t p0 = L.seg_pos[s], p1 = L.seg_pos[s + 1];
        for (uint32_t k = L.ut_off[s]; k < L.ut_off[s + 1]; ++k) {
            const uint32_t y = L.ut_tgt[k];
            if (!(y < L.n_top && stage[y] < stage[p0])) { err = "segment up target not an earlier top node"; return false; }
            for (uint32_t a = L.ut_aoff[k]; a < L.ut_aoff[k + 1]; ++a)
                if (L.uta_x[a] >= p1 - p0) { err = "segment up arc tail out of segment"; return false; }
        }
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
        std::vector<T3> want, got;
        const uint32_t nseg = (uint32_t)L.seg_pos.size() - 1;
        if (L.xr_off.size() != (size_t)nseg + 1 || L.pt_blk.size() != nseg + 1 ||
            L.tb_dn.size() != 64 * (size_t)nseg || L.tb_up.size() != 64 * (size_t)nseg ||
            L.xr_w.size() != 64 * (size_t)L.xr_src.size() || L.pr_w.size() != 64 * (size_t)L.pr_x.size() ||
            L.xr_off[nseg] != L.xr_src.size() || L.pr_off.size() != (size_t)L.pt_blk[nseg] + 1 ||
            L.pr_off.back() != L.pr_x.size() || L.pt_tgt.size() != 64 * (size_t)L.pt_blk[nseg]) {
            err = "dense tile tables disagree";
            return false;
        }
        for (uint32_t s = 0; s < nseg; ++s) {
            const uint32_t p0 = L.seg_pos[s], len = L.seg_pos[s + 1] - p0;
            if (len > 64) { err = "segment longer than a tile"; return false; }
            // DOWN from outside: (head offset, source, w)
            want.clear();
            got.clear();
            for (uint32_t i = 0; i < len; ++i)
                for (uint32_t k = L.in_off[p0 + i]; k < L.in_split[p0 + i]; ++k) want.emplace_back(i, L.in_src[k], L.in_w[k]);
            for (uint32_t r = L.xr_off[s]; r < L.xr_off[s + 1]; ++r) {
                const uint32_t u = L.xr_src[r];
                if (u >= n || stage[u] >= stage[p0]) { err = "tile source not final"; return false; }
                for (uint32_t j = 0; j < 64; ++j) {
                    const uint32_t w = L.xr_w[64 * (size_t)r + j];
                    if (w == INF) continue;
                    if (j >= len) { err = "tile head beyond the segment"; return false; }
                    got.emplace_back(j, u, w);
                }
            }
            merged(want);
            std::sort(got.begin(), got.end());
            if (want != got) { err = "DOWN tiles differ from the arcs they replace"; return false; }
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
                            err = "bad intra block";
                            return false;
                        }
                        for (uint32_t i = 0; i < 8; ++i)
                            for (uint32_t j = 0; j < 8; ++j) {
                                const uint32_t w = L.tb_w[64 * (size_t)k + 8 * i + j];
                                if (w == INF) continue;
                                const uint32_t t = 8 * a + i, h = 8 * o + j;
                                if (t >= len || h >= len || (dir == 0 ? t >= h : t <= h)) {
                                    err = "intra block arc against the sweep order";
                                    return false;
                                }
                                got.emplace_back(t, h, w);
                            }
                    }
                merged(want);
                std::sort(got.begin(), got.end());
                if (want != got) { err = "intra blocks differ from the arcs they replace"; return false; }
            }
            // UP leaving: (target, tail position, w)
            want.clear();
            got.clear();
            for (uint32_t i = 0; i < len; ++i)
                for (uint32_t k = L.up_off[p0 + i]; k < L.up_split[p0 + i]; ++k) want.emplace_back(L.up_tgt[k], p0 + i, L.up_w[k]);
            for (uint32_t kb = L.pt_blk[s]; kb < L.pt_blk[s + 1]; ++kb)
                for (uint32_t r = L.pr_off[kb]; r < L.pr_off[kb + 1]; ++r) {
                    const uint32_t x = L.pr_x[r];
                    if (x < p0 || x >= p0 + len) { err = "push tail outside the segment"; return false; }
                    for (uint32_t j = 0; j < 64; ++j) {
                        const uint32_t w = L.pr_w[64 * (size_t)r + j], t = L.pt_tgt[64 * (size_t)kb + j];
                        if (w == INF) continue;
                        if (t == NONE || t >= L.n_top || stage[t] >= stage[p0]) { err = "push target not an earlier top node"; return false; }
                        got.emplace_back(t, x, w);
                    }
                }
            merged(want);
            std::sort(got.begin(), got.end());
            if (want != got) { err = "UP push tiles differ from the arcs they replace"; return false; }
        }
        // DOWN work units: the units of a level cover the rows of its segments once
        if (L.du_lvl.size() != L.seglvl_seg.size() || L.du_r0.size() != L.du_seg.size() ||
            L.du_r1.size() != L.du_seg.size()) {
            err = "work unit tables disagree";
            return false;
        }
        std::vector<uint8_t> cover(L.xr_off[nseg], 0);
        for (size_t l = 0; l + 1 < L.du_lvl.size(); ++l)
            for (uint32_t u = L.du_lvl[l]; u < L.du_lvl[l + 1]; ++u) {
                const uint32_t s = L.du_seg[u];
                if (s >= nseg || lvl_of_seg[s] != l || L.du_r0[u] >= L.du_r1[u] || L.du_r0[u] < L.xr_off[s] ||
                    L.du_r1[u] > L.xr_off[s + 1]) {
                    err = "bad work unit";
                    return false;
                }
                for (uint32_t r = L.du_r0[u]; r < L.du_r1[u]; ++r) cover[r]++;
            }
        for (uint8_t c : cover)
            if (c != 1) { err = "work units do not cover the rows once"; return false; }
        if (L.pk_seg.size() != L.pt_blk[nseg]) { err = "target block table disagrees"; return false; }
        for (uint32_t s = 0; s < nseg; ++s)
            for (uint32_t k = L.pt_blk[s]; k < L.pt_blk[s + 1]; ++k)
                if (L.pk_seg[k] != s) { err = "target block of another segment"; return false; }
    }
    return true;
}

}  // namespace bfg
