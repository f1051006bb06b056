// layout.cpp — see layout.hpp.
#include "layout.hpp"

#include <algorithm>
#include <numeric>

namespace bfg {

namespace {
template <class T>
void prefix(std::vector<T> &v) {  // exclusive prefix sum in place, v has n+1 entries
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
    // ---- INF-filtered flats (bucket_ch.rs:191-424) ----
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
    L.up_off.assign(n + 1, 0);
    L.up_tgt.resize(mu);
    L.up_w.resize(mu);
    {
        // duplicate (tail, target) pairs keep the minimum weight, so the upward
        // walk relaxes each target once per tail
        uint32_t k = 0;
        std::vector<std::pair<uint32_t, uint32_t>> tmp;
        for (uint32_t r = 0; r < n; ++r) {
            L.up_off[r] = k;
            tmp.clear();
            for (uint64_t i = in.up_off[r]; i < in.up_off[r + 1]; ++i)
                if (in.up_w[i] != INF) tmp.emplace_back(in.up_tgt[i], in.up_w[i]);
            std::sort(tmp.begin(), tmp.end());
            for (size_t i = 0; i < tmp.size(); ++i) {
                if (i && tmp[i].first == tmp[i - 1].first) continue;
                L.up_tgt[k] = tmp[i].first; L.up_w[k] = tmp[i].second; ++k;
            }
        }
        L.up_off[n] = k;
        L.up_tgt.resize(k);
        L.up_w.resize(k);
    }
    // ---- elimination tree of the symmetrised pattern (Liu) ----
    // lower neighbours of j: down targets of j, and tails of up arcs into j
    std::vector<uint32_t> uin_off(n + 1, 0), uin_src(L.up_off[n]);
    for (uint32_t k = 0; k < L.up_off[n]; ++k) uin_off[L.up_tgt[k]]++;
    prefix(uin_off);
    {
        std::vector<uint32_t> cur(uin_off.begin(), uin_off.end() - 1);
        for (uint32_t r = 0; r < n; ++r)
            for (uint32_t k = L.up_off[r]; k < L.up_off[r + 1]; ++k) uin_src[cur[L.up_tgt[k]]++] = r;
    }
    L.parent.assign(n, NONE);
    {
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
    std::vector<uint32_t>().swap(uin_src);
    std::vector<uint32_t>().swap(uin_off);
    // depth, subtree size, preorder interval
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
    // ---- task schedule ----
    const uint32_t BL = std::max<uint32_t>(lp.bottom_limit, 1);
    std::vector<uint8_t> top(n, 0);
    std::vector<uint32_t> ntopc(n, 0);
    for (uint32_t v = 0; v < n; ++v) {
        top[v] = L.subtree[v] > BL;
        if (top[v] && L.parent[v] != NONE) ntopc[L.parent[v]]++;
    }
    // chains over top nodes, parents first
    std::vector<uint32_t> chain_of(n, NONE), chain_level;
    std::vector<std::vector<uint32_t>> chain_nodes;
    // bottom subtrees grouped by the chain of their parent (key = chain id + 1, 0 = no parent)
    std::vector<uint32_t> broot(n, NONE);
    std::vector<std::vector<uint32_t>> group_roots(1);  // index 0: parentless subtrees
    for (int64_t vv = (int64_t)n - 1; vv >= 0; --vv) {
        uint32_t v = (uint32_t)vv, p = L.parent[v];
        if (top[v]) {
            L.n_top++;
            if (p == NONE || ntopc[p] >= 2) {
                chain_of[v] = (uint32_t)chain_nodes.size();
                chain_level.push_back(p == NONE ? 0 : chain_level[chain_of[p]] + 1);
                chain_nodes.emplace_back();
                group_roots.emplace_back();
            } else {
                chain_of[v] = chain_of[p];
            }
            chain_nodes[chain_of[v]].push_back(v);
        } else {
            if (p == NONE || top[p]) {
                broot[v] = v;
                group_roots[p == NONE ? 0 : chain_of[p] + 1].push_back(v);
            } else {
                broot[v] = broot[p];
            }
        }
    }
    L.n_chains = (uint32_t)chain_nodes.size();
    // bottom tasks: pack subtrees of one group up to group_limit nodes
    struct BTask { uint32_t level; std::vector<uint32_t> roots; };
    std::vector<BTask> btasks;
    std::vector<uint32_t> btask_of_root(n, NONE);
    const uint32_t GL = std::max(lp.group_limit, BL);
    for (size_t gi = 0; gi < group_roots.size(); ++gi) {
        auto &roots = group_roots[gi];
        if (roots.empty()) continue;
        uint32_t lvl = gi == 0 ? 0 : chain_level[gi - 1] + 1;
        uint32_t acc = 0;
        bool fresh = true;
        for (uint32_t r : roots) {
            if (fresh || acc + L.subtree[r] > GL) { btasks.push_back({lvl, {}}); acc = 0; fresh = false; }
            btasks.back().roots.push_back(r);
            btask_of_root[r] = (uint32_t)btasks.size() - 1;
            acc += L.subtree[r];
        }
    }
    L.n_bottom_tasks = (uint32_t)btasks.size();
    // nodes of each bottom task bucketed by relative depth
    std::vector<std::vector<std::vector<uint32_t>>> bsteps(btasks.size());
    for (uint32_t v = 0; v < n; ++v) {
        if (top[v]) continue;
        uint32_t r = broot[v];
        uint32_t t = btask_of_root[r];
        uint32_t rel = L.depth[v] - L.depth[r];
        auto &st = bsteps[t];
        if (st.size() <= rel) st.resize(rel + 1);
        st[rel].push_back(v);  // ascending rank inside a step
    }
    // ---- positions: level-major ----
    uint32_t max_level = 0;
    for (uint32_t l : chain_level) max_level = std::max(max_level, l);
    for (auto &b : btasks) max_level = std::max(max_level, b.level);
    std::vector<std::vector<uint32_t>> lvl_chains(max_level + 1), lvl_btasks(max_level + 1);
    for (uint32_t c = 0; c < chain_nodes.size(); ++c) lvl_chains[chain_level[c]].push_back(c);
    for (uint32_t t = 0; t < btasks.size(); ++t) lvl_btasks[btasks[t].level].push_back(t);
    L.pos_of_rank.assign(n, NONE);
    L.rank_of_pos.assign(n, NONE);
    uint32_t pos = 0;
    L.step_pos.clear(); L.task_step.clear(); L.level_task.clear();
    for (uint32_t l = 0; l <= max_level; ++l) {
        L.level_task.push_back((uint32_t)L.task_step.size());
        for (uint32_t c : lvl_chains[l]) {
            L.task_step.push_back((uint32_t)L.step_pos.size());
            for (uint32_t v : chain_nodes[c]) {
                L.step_pos.push_back(pos);
                L.pos_of_rank[v] = pos; L.rank_of_pos[pos] = v; ++pos;
            }
            L.max_steps_per_task = std::max<uint32_t>(L.max_steps_per_task, (uint32_t)chain_nodes[c].size());
        }
        for (uint32_t t : lvl_btasks[l]) {
            L.task_step.push_back((uint32_t)L.step_pos.size());
            for (auto &st : bsteps[t]) {
                L.step_pos.push_back(pos);
                for (uint32_t v : st) { L.pos_of_rank[v] = pos; L.rank_of_pos[pos] = v; ++pos; }
            }
            L.max_steps_per_task = std::max<uint32_t>(L.max_steps_per_task, (uint32_t)bsteps[t].size());
        }
    }
    L.level_task.push_back((uint32_t)L.task_step.size());
    L.task_step.push_back((uint32_t)L.step_pos.size());
    L.step_pos.push_back(pos);
    if (pos != n) { err = "layout lost nodes"; return false; }
    // ---- pull flat by position (DownReverseAdjFlat, bucket_ch.rs:343-424) ----
    L.in_off.assign(n + 1, 0);
    for (uint32_t r = 0; r < n; ++r)
        for (uint64_t i = in.dn_off[r]; i < in.dn_off[r + 1]; ++i)
            if (in.dn_w[i] != INF) L.in_off[L.pos_of_rank[in.dn_tgt[i]]]++;
    prefix(L.in_off);
    L.in_src.resize(md);
    L.in_w.resize(md);
    {
        std::vector<uint32_t> cur(L.in_off.begin(), L.in_off.end() - 1);
        for (uint32_t r = 0; r < n; ++r) {
            uint32_t pr = L.pos_of_rank[r];
            for (uint64_t i = in.dn_off[r]; i < in.dn_off[r + 1]; ++i) {
                if (in.dn_w[i] == INF) continue;
                uint32_t p = L.pos_of_rank[in.dn_tgt[i]];
                uint32_t k = cur[p]++;
                L.in_src[k] = pr;
                L.in_w[k] = in.dn_w[i];
            }
        }
        // sort each node's in-arcs by source position
        std::vector<std::pair<uint32_t, uint32_t>> tmp;
        for (uint32_t p = 0; p < n; ++p) {
            uint32_t a = L.in_off[p], b = L.in_off[p + 1];
            if (b - a < 2) continue;
            tmp.resize(b - a);
            for (uint32_t k = a; k < b; ++k) tmp[k - a] = {L.in_src[k], L.in_w[k]};
            std::sort(tmp.begin(), tmp.end());
            for (uint32_t k = a; k < b; ++k) { L.in_src[k] = tmp[k - a].first; L.in_w[k] = tmp[k - a].second; }
        }
    }
    L.pos_pre.resize(n); L.pos_subtree.resize(n); L.pos_depth.resize(n);
    for (uint32_t p = 0; p < n; ++p) {
        uint32_t v = L.rank_of_pos[p];
        L.pos_pre[p] = L.pre[v];
        L.pos_subtree[p] = L.subtree[v];
        L.pos_depth[p] = L.depth[v];
    }
    return true;
}

bool check_ancestry(const CchInput &in, const Layout &L, std::string &err) {
    auto is_anc = [&](uint32_t a, uint32_t d) {  // a ancestor-or-self of d
        return L.pre[a] <= L.pre[d] && L.pre[d] < L.pre[a] + L.subtree[a];
    };
    for (uint32_t r = 0; r < in.n; ++r) {
        for (uint64_t i = in.up_off[r]; i < in.up_off[r + 1]; ++i)
            if (in.up_w[i] != INF && !is_anc(in.up_tgt[i], r)) { err = "up arc between non-ancestors"; return false; }
        for (uint64_t i = in.dn_off[r]; i < in.dn_off[r + 1]; ++i)
            if (in.dn_w[i] != INF && !is_anc(r, in.dn_tgt[i])) { err = "down arc between non-ancestors"; return false; }
    }
    // schedule: every in-arc source is final before its head's step starts —
    // an earlier step of the same task, or a task of an earlier level.
    const uint32_t n = in.n;
    std::vector<uint32_t> step_of(n), task_of_step(L.step_pos.size() - 1), level_of_task(L.task_step.size() - 1);
    for (size_t l = 0; l + 1 < L.level_task.size(); ++l)
        for (uint32_t t = L.level_task[l]; t < L.level_task[l + 1]; ++t) level_of_task[t] = (uint32_t)l;
    for (size_t t = 0; t + 1 < L.task_step.size(); ++t)
        for (uint32_t s = L.task_step[t]; s < L.task_step[t + 1]; ++s) task_of_step[s] = (uint32_t)t;
    for (size_t s = 0; s + 1 < L.step_pos.size(); ++s)
        for (uint32_t p = L.step_pos[s]; p < L.step_pos[s + 1]; ++p) step_of[p] = (uint32_t)s;
    for (uint32_t p = 0; p < n; ++p) {
        uint32_t sp = step_of[p], tp = task_of_step[sp];
        for (uint32_t k = L.in_off[p]; k < L.in_off[p + 1]; ++k) {
            uint32_t q = L.in_src[k], sq = step_of[q], tq = task_of_step[sq];
            bool ok = (tq == tp) ? (sq < sp) : (level_of_task[tq] < level_of_task[tp]);
            if (!ok) { err = "schedule violates an arc dependency"; return false; }
        }
    }
    return true;
}

}  // namespace bfg
