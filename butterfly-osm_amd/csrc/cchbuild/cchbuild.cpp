// cchbuild.cpp — seeded synthetic road grid -> customized, rank-aligned
// edge-based CCH.  Host C++17 + OpenMP.  See cchbuild.h for the reference
// steps each stage restates.
#include "cchbuild.h"

#include <omp.h>

#include <algorithm>
#include <atomic>
#include <memory>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <numeric>
#include <queue>
#include <string>
#include <tuple>
#include <vector>

namespace {

thread_local std::string g_err;
constexpr uint32_t INF = 0xFFFFFFFFu;

inline uint32_t sat_add(uint32_t a, uint32_t b) {
    uint32_t s = a + b;
    return s < a ? INF : s;
}

inline uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// xoshiro256** seeded through splitmix64.
struct Rng {
    uint64_t s[4];
    explicit Rng(uint64_t seed) {
        uint64_t z = seed;
        for (auto &w : s) { z = mix64(z); w = z; }
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next() {
        uint64_t r = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
        s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
        s[2] ^= t; s[3] = rotl(s[3], 45);
        return r;
    }
    double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }
    uint64_t below(uint64_t n) { return next() % n; }
};

double now_s() {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

// route/src/weights.rs:14-40
uint64_t round_half_even_div(uint64_t num, uint64_t den) {
    if (den == 0) return 0;
    uint64_t q = num / den, r = num % den;
    uint64_t twice = r * 2;
    bool of = r > (UINT64_MAX >> 1);
    if (!of && twice < den) return q;
    if (of || twice > den) return q + 1;
    return (q % 2 == 0) ? q : q + 1;
}

// route/src/model/compile.rs:141-142
uint32_t kmh_to_mmps(double kmh) { return (uint32_t)std::llround(kmh * 1000.0 / 3.6); }

// Highway classes of the synthetic grid -> car-model speeds
// (models/car.model.json: tertiary_link 30, tertiary 50, primary 90, motorway 120)
const double kClassKmh[4] = {30.0, 50.0, 90.0, 120.0};

int class_of_index(int idx) {
    if (idx % 32 == 0) return 3;
    if (idx % 8 == 0) return 2;
    if (idx % 2 == 0) return 1;
    return 0;
}

struct Edge {
    uint32_t a, b;
    uint8_t cls;
    uint8_t oneway;  // 0 both, 1 a->b only, 2 b->a only
    uint32_t length_m;
};

struct Seg {
    uint32_t tail, head, edge;
    uint32_t weight_s;
    uint16_t bearing_deci;
};

// route/src/ebg/turn_penalty.rs:234-288 with TurnPenaltyConfig::car()
uint32_t car_turn_penalty(uint16_t from_b, uint16_t to_b, uint32_t via_degree) {
    const uint32_t turn_penalty_s = 8, u_turn_penalty_s = 20, min_degree = 3;
    const double turn_bias = 1.075;
    if (via_degree < min_degree) return 0;
    double delta = to_b / 10.0 - from_b / 10.0;
    while (delta > 180.0) delta -= 360.0;
    while (delta < -180.0) delta += 360.0;
    int angle = (int)std::round(delta);
    bool is_uturn = std::abs(angle) >= 170;
    double exponent = -((13.0 / turn_bias) * (-(double)angle / 180.0) - 6.5 * turn_bias);
    double sigmoid = 1.0 / (1.0 + std::exp(exponent));
    uint32_t pen = (uint32_t)std::round(turn_penalty_s * sigmoid);
    if (is_uturn) pen = sat_add(pen, u_turn_penalty_s);
    return pen;
}

uint16_t bearing_deci(double dx, double dy) {
    double deg = std::atan2(dx, dy) * 180.0 / M_PI;  // 0 = north, clockwise
    if (deg < 0) deg += 360.0;
    long d = std::lround(deg * 10.0) % 3600;
    return (uint16_t)d;
}

// ---------------------------------------------------------------------------
// Graph: NBG -> segments -> EBG (CSR by segment id)
// ---------------------------------------------------------------------------
struct Ebg {
    uint32_t n = 0;
    std::vector<uint64_t> off;
    std::vector<uint32_t> head, w;
    std::vector<float> x, y;
    uint32_t nbg_nodes = 0, nbg_edges = 0;
};

bool build_graph(const bfc_params &p, Ebg &g) {
    const int R = p.rows, C = p.cols;
    if (R < 2 || C < 2) { g_err = "grid must be at least 2x2"; return false; }
    Rng rng(p.seed);
    const uint32_t nn = (uint32_t)R * (uint32_t)C;
    std::vector<double> nx(nn), ny(nn);
    const double jit = p.jitter_frac * p.spacing_m;
    for (int r = 0; r < R; ++r)
        for (int c = 0; c < C; ++c) {
            uint32_t id = (uint32_t)r * C + c;
            nx[id] = c * p.spacing_m + (2.0 * rng.uniform() - 1.0) * jit;
            ny[id] = r * p.spacing_m + (2.0 * rng.uniform() - 1.0) * jit;
        }
    std::vector<Edge> edges;
    edges.reserve((size_t)nn * 2 + 16);
    auto add_edge = [&](uint32_t a, uint32_t b, int cls) {
        Edge e;
        e.a = a; e.b = b; e.cls = (uint8_t)cls;
        e.oneway = 0;
        if (rng.uniform() < p.oneway_frac) e.oneway = (rng.uniform() < 0.5) ? 1 : 2;
        double dx = nx[b] - nx[a], dy = ny[b] - ny[a];
        double len = std::sqrt(dx * dx + dy * dy);
        e.length_m = (uint32_t)std::max(1.0, std::round(len));
        edges.push_back(e);
    };
    for (int r = 0; r < R; ++r)
        for (int c = 0; c < C; ++c) {
            uint32_t id = (uint32_t)r * C + c;
            if (c + 1 < C) {
                if (rng.uniform() >= p.delete_frac) add_edge(id, id + 1, class_of_index(r));
            }
            if (r + 1 < R) {
                if (rng.uniform() >= p.delete_frac) add_edge(id, id + C, class_of_index(c));
            }
        }
    // long motorway chords: straight links 4..32 cells long in a random direction
    uint64_t n_chords = (uint64_t)std::llround(p.chord_frac * nn);
    for (uint64_t i = 0; i < n_chords; ++i) {
        uint32_t a = (uint32_t)rng.below(nn);
        double L = 4.0 + 28.0 * rng.uniform();
        double th = 2.0 * M_PI * rng.uniform();
        int ra = (int)(a / C), ca = (int)(a % C);
        int rb = std::min(R - 1, std::max(0, ra + (int)std::lround(L * std::sin(th))));
        int cb = std::min(C - 1, std::max(0, ca + (int)std::lround(L * std::cos(th))));
        uint32_t b = (uint32_t)rb * C + cb;
        if (a != b) add_edge(a, b, 3);
    }
    g.nbg_nodes = nn;
    g.nbg_edges = (uint32_t)edges.size();

    // segments (EBG nodes): forward then reverse per edge
    std::vector<Seg> segs;
    segs.reserve(edges.size() * 2);
    std::vector<uint32_t> degree(nn, 0);
    for (uint32_t ei = 0; ei < edges.size(); ++ei) {
        const Edge &e = edges[ei];
        degree[e.a]++; degree[e.b]++;
        uint32_t mmps = kmh_to_mmps(kClassKmh[e.cls]);
        uint32_t w = (uint32_t)round_half_even_div((uint64_t)e.length_m * 1000, mmps);
        w = std::max<uint32_t>(w, 1);
        double dx = nx[e.b] - nx[e.a], dy = ny[e.b] - ny[e.a];
        if (e.oneway != 2) segs.push_back({e.a, e.b, ei, w, bearing_deci(dx, dy)});
        if (e.oneway != 1) segs.push_back({e.b, e.a, ei, w, bearing_deci(-dx, -dy)});
    }
    const uint32_t ns = (uint32_t)segs.size();
    // out segments per NBG node (CSR)
    std::vector<uint32_t> out_off(nn + 1, 0), out_seg(ns);
    for (auto &s : segs) out_off[s.tail + 1]++;
    for (uint32_t i = 0; i < nn; ++i) out_off[i + 1] += out_off[i];
    {
        std::vector<uint32_t> cur(out_off.begin(), out_off.end() - 1);
        for (uint32_t si = 0; si < ns; ++si) out_seg[cur[segs[si].tail]++] = si;
    }
    // turn arcs (route/src/ebg/mod.rs:370-440)
    const uint64_t ban_key = mix64(p.seed ^ 0x7475726e62616e73ull);
    g.n = ns;
    g.off.assign(ns + 1, 0);
    std::vector<uint32_t> cnt(ns, 0);
    auto emit = [&](uint32_t a, bool write, uint64_t pos) -> uint32_t {
        const Seg &sa = segs[a];
        uint32_t v = sa.head;
        uint32_t n_out = out_off[v + 1] - out_off[v];
        uint32_t k = 0;
        for (uint32_t j = out_off[v]; j < out_off[v + 1]; ++j) {
            uint32_t b = out_seg[j];
            const Seg &sb = segs[b];
            bool uturn = sb.head == sa.tail;
            if (uturn && n_out != 1) continue;
            if (!uturn) {
                uint64_t h = mix64(ban_key ^ ((uint64_t)a << 32 | b));
                if ((double)(h >> 11) * 0x1.0p-53 < p.ban_frac) continue;
            }
            if (write) {
                uint32_t pen = car_turn_penalty(sa.bearing_deci, sb.bearing_deci,
                                                std::min<uint32_t>(degree[v], 255));
                g.head[pos + k] = b;
                g.w[pos + k] = sat_add(sb.weight_s, pen);
            }
            ++k;
        }
        return k;
    };
#pragma omp parallel for schedule(static)
    for (int64_t a = 0; a < (int64_t)ns; ++a) cnt[a] = emit((uint32_t)a, false, 0);
    for (uint32_t a = 0; a < ns; ++a) g.off[a + 1] = g.off[a] + cnt[a];
    g.head.resize(g.off[ns]);
    g.w.resize(g.off[ns]);
#pragma omp parallel for schedule(static)
    for (int64_t a = 0; a < (int64_t)ns; ++a) emit((uint32_t)a, true, g.off[a]);
    g.x.resize(ns); g.y.resize(ns);
    for (uint32_t si = 0; si < ns; ++si) {
        g.x[si] = (float)(0.5 * (nx[segs[si].tail] + nx[segs[si].head]));
        g.y[si] = (float)(0.5 * (ny[segs[si].tail] + ny[segs[si].head]));
    }
    return true;
}

// ---------------------------------------------------------------------------
// Ordering (route/src/ordering.rs)
// ---------------------------------------------------------------------------
struct Orderer {
    const Ebg &g;
    int leaf;
    std::vector<int32_t> loc;   // global -> local id scratch (disjoint per call)
    std::vector<int8_t> side;   // 0 none, 1 A, 2 B (disjoint per call)
    std::vector<uint32_t> rin_off, rin_src;  // reverse adjacency for components
    explicit Orderer(const Ebg &g_, int leaf_) : g(g_), leaf(leaf_) {
        loc.assign(g.n, -1);
        side.assign(g.n, 0);
    }

    // ordering.rs:359-431 — weakly connected components, sorted by size desc then min id
    std::vector<std::vector<uint32_t>> components() {
        const uint32_t n = g.n;
        rin_off.assign(n + 1, 0);
        for (uint64_t i = 0; i < g.head.size(); ++i) rin_off[g.head[i] + 1]++;
        for (uint32_t i = 0; i < n; ++i) rin_off[i + 1] += rin_off[i];
        rin_src.resize(g.head.size());
        {
            std::vector<uint32_t> cur(rin_off.begin(), rin_off.end() - 1);
            for (uint32_t u = 0; u < n; ++u)
                for (uint64_t i = g.off[u]; i < g.off[u + 1]; ++i) rin_src[cur[g.head[i]]++] = u;
        }
        std::vector<char> vis(n, 0);
        std::vector<std::vector<uint32_t>> comps;
        std::deque<uint32_t> q;
        for (uint32_t s = 0; s < n; ++s) {
            if (vis[s]) continue;
            std::vector<uint32_t> comp;
            vis[s] = 1; q.push_back(s);
            while (!q.empty()) {
                uint32_t u = q.front(); q.pop_front();
                comp.push_back(u);
                for (uint64_t i = g.off[u]; i < g.off[u + 1]; ++i) {
                    uint32_t v = g.head[i];
                    if (!vis[v]) { vis[v] = 1; q.push_back(v); }
                }
                for (uint32_t i = rin_off[u]; i < rin_off[u + 1]; ++i) {
                    uint32_t v = rin_src[i];
                    if (!vis[v]) { vis[v] = 1; q.push_back(v); }
                }
            }
            comps.push_back(std::move(comp));
        }
        std::stable_sort(comps.begin(), comps.end(), [](const auto &a, const auto &b) {
            if (a.size() != b.size()) return a.size() > b.size();
            return *std::min_element(a.begin(), a.end()) < *std::min_element(b.begin(), b.end());
        });
        return comps;
    }

    // ordering.rs:69-176 — minimum degree with (degree, global id) tie-break
    std::vector<uint32_t> min_degree(const std::vector<uint32_t> &nodes) {
        const size_t n = nodes.size();
        std::vector<uint32_t> out;
        if (n == 0) return out;
        for (size_t i = 0; i < n; ++i) loc[nodes[i]] = (int32_t)i;
        std::vector<std::vector<uint32_t>> adj(n);
        for (size_t u = 0; u < n; ++u) {
            uint32_t gu = nodes[u];
            for (uint64_t i = g.off[gu]; i < g.off[gu + 1]; ++i) {
                int32_t v = loc[g.head[i]];
                if (v >= 0 && (size_t)v != u) {
                    adj[u].push_back((uint32_t)v);
                    adj[v].push_back((uint32_t)u);
                }
            }
        }
        for (auto &l : adj) { std::sort(l.begin(), l.end()); l.erase(std::unique(l.begin(), l.end()), l.end()); }
        std::vector<size_t> deg(n);
        for (size_t u = 0; u < n; ++u) deg[u] = adj[u].size();
        std::vector<char> elim(n, 0);
        typedef std::tuple<size_t, uint32_t, uint32_t> E;
        std::priority_queue<E, std::vector<E>, std::greater<E>> heap;
        for (size_t u = 0; u < n; ++u) heap.emplace(deg[u], nodes[u], (uint32_t)u);
        out.reserve(n);
        std::vector<uint32_t> nb;
        for (size_t it = 0; it < n; ++it) {
            uint32_t m;
            for (;;) {
                E e = heap.top(); heap.pop();
                uint32_t u = std::get<2>(e);
                if (!elim[u] && deg[u] == std::get<0>(e)) { m = u; break; }
            }
            elim[m] = 1;
            out.push_back(nodes[m]);
            nb.clear();
            for (uint32_t v : adj[m]) if (!elim[v]) nb.push_back(v);
            for (size_t i = 0; i < nb.size(); ++i)
                for (size_t j = i + 1; j < nb.size(); ++j) {
                    uint32_t u = nb[i], v = nb[j];
                    auto it_u = std::lower_bound(adj[u].begin(), adj[u].end(), v);
                    if (it_u == adj[u].end() || *it_u != v) {
                        adj[u].insert(it_u, v);
                        auto it_v = std::lower_bound(adj[v].begin(), adj[v].end(), u);
                        adj[v].insert(it_v, u);
                        deg[u]++; deg[v]++;
                        heap.emplace(deg[u], nodes[u], u);
                        heap.emplace(deg[v], nodes[v], v);
                    }
                }
            for (uint32_t v : nb) {
                auto it_v = std::lower_bound(adj[v].begin(), adj[v].end(), m);
                if (it_v != adj[v].end() && *it_v == m) {
                    adj[v].erase(it_v);
                    if (deg[v] > 0) deg[v]--;
                    heap.emplace(deg[v], nodes[v], v);
                }
            }
        }
        for (size_t i = 0; i < n; ++i) loc[nodes[i]] = -1;
        return out;
    }

    // ordering.rs:715-774
    static void histogram_partition(const std::vector<std::pair<double, uint32_t>> &proj,
                                    std::vector<uint32_t> &A, std::vector<uint32_t> &B) {
        const size_t NB = 4096;
        if (proj.empty()) return;
        double lo = INFINITY, hi = -INFINITY;
        for (auto &pr : proj) { lo = std::min(lo, pr.first); hi = std::max(hi, pr.first); }
        double range = std::max(hi - lo, 1e-9);
        std::vector<std::vector<uint32_t>> hist(NB);
        for (auto &pr : proj) {
            double t = std::min(std::max((pr.first - lo) / range, 0.0), 0.9999);
            hist[(size_t)(t * (double)NB)].push_back(pr.second);
        }
        for (auto &b : hist) std::sort(b.begin(), b.end());
        size_t target = proj.size() / 2, cum = 0, median = 0;
        for (size_t i = 0; i < NB; ++i) {
            cum += hist[i].size();
            if (cum >= target) { median = i; break; }
        }
        for (size_t i = 0; i < NB; ++i) {
            if (i < median) A.insert(A.end(), hist[i].begin(), hist[i].end());
            else if (i > median) B.insert(B.end(), hist[i].begin(), hist[i].end());
            else {
                size_t need = target > A.size() ? target - A.size() : 0;
                size_t k = std::min(need, hist[i].size());
                A.insert(A.end(), hist[i].begin(), hist[i].begin() + k);
                B.insert(B.end(), hist[i].begin() + k, hist[i].end());
            }
        }
    }

    // ordering.rs:591-680 — greedy vertex cover of the A<->B cross arcs
    std::vector<uint32_t> separator(const std::vector<uint32_t> &A, const std::vector<uint32_t> &B) {
        for (uint32_t u : A) side[u] = 1;
        for (uint32_t u : B) side[u] = 2;
        std::vector<std::pair<uint32_t, uint32_t>> cross;
        for (uint32_t u : A)
            for (uint64_t i = g.off[u]; i < g.off[u + 1]; ++i)
                if (side[g.head[i]] == 2) cross.emplace_back(u, g.head[i]);
        for (uint32_t u : B)
            for (uint64_t i = g.off[u]; i < g.off[u + 1]; ++i)
                if (side[g.head[i]] == 1) cross.emplace_back(u, g.head[i]);
        for (uint32_t u : A) side[u] = 0;
        for (uint32_t u : B) side[u] = 0;
        std::vector<uint32_t> sep;
        if (cross.empty()) return sep;
        // node -> incident cross-edge indices
        std::vector<uint32_t> ends;
        ends.reserve(cross.size() * 2);
        for (auto &e : cross) { ends.push_back(e.first); ends.push_back(e.second); }
        std::sort(ends.begin(), ends.end());
        ends.erase(std::unique(ends.begin(), ends.end()), ends.end());
        auto idx = [&](uint32_t v) { return (size_t)(std::lower_bound(ends.begin(), ends.end(), v) - ends.begin()); };
        std::vector<uint32_t> eoff(ends.size() + 1, 0), elist(cross.size() * 2);
        for (auto &e : cross) { eoff[idx(e.first) + 1]++; eoff[idx(e.second) + 1]++; }
        for (size_t i = 0; i < ends.size(); ++i) eoff[i + 1] += eoff[i];
        {
            std::vector<uint32_t> cur(eoff.begin(), eoff.end() - 1);
            for (uint32_t k = 0; k < cross.size(); ++k) {
                elist[cur[idx(cross[k].first)]++] = k;
                elist[cur[idx(cross[k].second)]++] = k;
            }
        }
        std::vector<uint32_t> order(ends.size());
        std::iota(order.begin(), order.end(), 0);
        std::sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
            uint32_t da = eoff[a + 1] - eoff[a], db = eoff[b + 1] - eoff[b];
            if (da != db) return da > db;
            return ends[a] < ends[b];
        });
        std::vector<char> covered(cross.size(), 0);
        size_t n_cov = 0;
        for (uint32_t oi : order) {
            if (n_cov == cross.size()) break;
            bool fresh = false;
            for (uint32_t k = eoff[oi]; k < eoff[oi + 1]; ++k) if (!covered[elist[k]]) { fresh = true; break; }
            if (fresh) {
                sep.push_back(ends[oi]);
                for (uint32_t k = eoff[oi]; k < eoff[oi + 1]; ++k)
                    if (!covered[elist[k]]) { covered[elist[k]] = 1; ++n_cov; }
            }
        }
        std::sort(sep.begin(), sep.end());
        return sep;
    }

    // ordering.rs:480-589
    std::vector<uint32_t> nd(const std::vector<uint32_t> &nodes) {
        if ((int64_t)nodes.size() <= leaf) return min_degree(nodes);
        if (nodes.size() <= 2) return min_degree(nodes);
        double mx = 0, my = 0;
        for (uint32_t u : nodes) { mx += g.x[u]; my += g.y[u]; }
        mx /= (double)nodes.size(); my /= (double)nodes.size();
        double cxx = 0, cxy = 0, cyy = 0;
        for (uint32_t u : nodes) {
            double dx = g.x[u] - mx, dy = g.y[u] - my;
            cxx += dx * dx; cxy += dx * dy; cyy += dy * dy;
        }
        // ordering.rs:693-712
        double tr = cxx + cyy, det = cxx * cyy - cxy * cxy;
        double disc = std::max(tr * tr / 4.0 - det, 0.0);
        double l1 = tr / 2.0 + std::sqrt(disc);
        double dirx = 1.0, diry = 0.0;
        if (std::fabs(cxy) > 1e-10) {
            double vx = l1 - cyy, vy = cxy, nrm = std::sqrt(vx * vx + vy * vy);
            if (nrm > 1e-10) { dirx = vx / nrm; diry = vy / nrm; }
        }
        std::vector<std::pair<double, uint32_t>> proj;
        proj.reserve(nodes.size());
        for (uint32_t u : nodes) proj.emplace_back((g.x[u] - mx) * dirx + (g.y[u] - my) * diry, u);
        std::vector<uint32_t> A, B;
        histogram_partition(proj, A, B);
        std::vector<uint32_t> sep = separator(A, B);
        if (!sep.empty()) {
            for (uint32_t u : sep) side[u] = 3;
            auto keep = [&](std::vector<uint32_t> &v) {
                v.erase(std::remove_if(v.begin(), v.end(), [&](uint32_t u) { return side[u] == 3; }), v.end());
            };
            keep(A); keep(B);
            for (uint32_t u : sep) side[u] = 0;
        }
        double bal = (double)A.size() / (double)std::max<size_t>(A.size() + B.size(), 1);
        if (!(bal >= 0.2 && bal <= 0.8)) return min_degree(nodes);
        std::vector<uint32_t> oa, ob;
        const size_t PAR = 50000;
        if (A.size() >= PAR && B.size() >= PAR) {
#pragma omp task shared(oa, A)
            oa = nd(A);
#pragma omp task shared(ob, B)
            ob = nd(B);
#pragma omp taskwait
        } else {
            oa = nd(A);
            ob = nd(B);
        }
        oa.insert(oa.end(), ob.begin(), ob.end());
        oa.insert(oa.end(), sep.begin(), sep.end());
        return oa;
    }
};

// ---------------------------------------------------------------------------
// Contraction + basic customization in one bottom-up pass.
// An arc u->w is kept at its lower endpoint: out_hi[u] if u < w, in_hi[w] if
// u > w.  When x is contracted every arc at x is final, and each pair
// (u in in_hi[x], w in out_hi[x]) yields arc u->w with weight d(u,x)+d(x,w),
// merged (min) into the lower endpoint's list — the in x out fill of
// route/src/contraction.rs:247-400 together with the lower-triangle relaxation
// of route/src/customization.rs:1005-1200.
// ---------------------------------------------------------------------------
struct Ent { uint32_t v, w; };
typedef std::vector<Ent> List;

// Merges the sorted entries add[0..na) — key add[j].v, weight base + add[j].w
// (saturating) — into sorted `dst`, keeping the minimum weight on equal keys.
// Returns the number of keys that were new.  The common clique case (every key
// already present) is updated in place without allocation.
size_t merge_add(List &dst, const Ent *add, size_t na, uint32_t base, List &tmp) {
    const size_t nd = dst.size();
    Ent *d = dst.data();
    size_t fresh = 0;
    if (na * 16 < nd) {
        size_t lo = 0;
        for (size_t j = 0; j < na; ++j) {
            const uint32_t key = add[j].v;
            size_t pos = (size_t)(std::lower_bound(d + lo, d + nd, key,
                                                   [](const Ent &e, uint32_t k) { return e.v < k; }) - d);
            if (pos < nd && d[pos].v == key) {
                const uint32_t cand = sat_add(base, add[j].w);
                if (cand < d[pos].w) d[pos].w = cand;
                lo = pos + 1;
            } else {
                ++fresh;
                lo = pos;
            }
        }
    } else {
        size_t i = 0, j = 0;
        while (i < nd && j < na) {
            const uint32_t a = d[i].v, b = add[j].v;
            if (a == b) {
                const uint32_t cand = sat_add(base, add[j].w);
                d[i].w = cand < d[i].w ? cand : d[i].w;
            }
            fresh += (b < a);
            i += (a <= b);
            j += (b <= a);
        }
        fresh += na - j;
    }
    if (fresh == 0) return 0;
    tmp.resize(nd + fresh);
    Ent *o = tmp.data();
    size_t i = 0, j = 0, k = 0;
    while (i < nd && j < na) {
        if (d[i].v < add[j].v) o[k++] = d[i++];
        else if (d[i].v > add[j].v) { o[k++] = {add[j].v, sat_add(base, add[j].w)}; ++j; }
        else { o[k++] = d[i]; ++i; ++j; }
    }
    while (i < nd) o[k++] = d[i++];
    while (j < na) { o[k++] = {add[j].v, sat_add(base, add[j].w)}; ++j; }
    tmp.resize(k);
    dst.swap(tmp);
    return fresh;
}

}  // namespace

extern "C" {

const char *bfc_last_error(void) { return g_err.c_str(); }

void bfc_free(bfc_cch *c) {
    if (!c) return;
    free(c->up_off); free(c->up_tgt); free(c->up_w);
    free(c->dn_off); free(c->dn_tgt); free(c->dn_w);
    free(c->x); free(c->y); free(c->rank_to_filtered);
    free(c->ebg_off); free(c->ebg_head); free(c->ebg_w);
    free(c);
}

bfc_cch *bfc_build_grid(const bfc_params *pp) {
    if (!pp) { g_err = "null params"; return nullptr; }
    bfc_params p = *pp;
    if (p.leaf_threshold <= 0) p.leaf_threshold = 8192;
    if (p.n_threads > 0) omp_set_num_threads(p.n_threads);
    const bool vb = p.verbose != 0;
    double t0 = now_s();
    Ebg g;
    if (!build_graph(p, g)) return nullptr;
    double t1 = now_s();
    if (vb) fprintf(stderr, "[cchbuild] graph: nbg %u nodes %u edges -> ebg %u nodes %zu arcs (%.2fs)\n",
                    g.nbg_nodes, g.nbg_edges, g.n, g.head.size(), t1 - t1 + (t1 - t0));
    const uint32_t n = g.n;
    // ---- ordering ----
    Orderer ord(g, p.leaf_threshold);
    auto comps = ord.components();
    std::vector<uint32_t> inv_perm;  // rank -> filtered id
    inv_perm.reserve(n);
    for (auto &comp : comps) {
        std::vector<uint32_t> o;
#pragma omp parallel
#pragma omp single
        o = ord.nd(comp);
        inv_perm.insert(inv_perm.end(), o.begin(), o.end());
    }
    if (inv_perm.size() != n) { g_err = "ordering lost nodes"; return nullptr; }
    std::vector<uint32_t> perm(n, INF);
    for (uint32_t r = 0; r < n; ++r) perm[inv_perm[r]] = r;
    for (uint32_t v = 0; v < n; ++v) if (perm[v] == INF) { g_err = "ordering not a permutation"; return nullptr; }
    double t2 = now_s();
    if (vb) fprintf(stderr, "[cchbuild] order: %zu components (%.2fs)\n", comps.size(), t2 - t1);

    // ---- contraction + customization (rank space) ----
    std::vector<List> out_hi(n), in_hi(n);
    for (uint32_t u = 0; u < n; ++u) {
        uint32_t ru = perm[u];
        for (uint64_t i = g.off[u]; i < g.off[u + 1]; ++i) {
            uint32_t rv = perm[g.head[i]];
            if (rv == ru) continue;
            if (rv > ru) out_hi[ru].push_back({rv, g.w[i]});
            else in_hi[rv].push_back({ru, g.w[i]});
        }
    }
    auto norm = [](List &l) {
        std::sort(l.begin(), l.end(), [](const Ent &a, const Ent &b) { return a.v < b.v || (a.v == b.v && a.w < b.w); });
        size_t k = 0;
        for (size_t i = 0; i < l.size(); ++i)
            if (k == 0 || l[k - 1].v != l[i].v) l[k++] = l[i];
        l.resize(k);
    };
#pragma omp parallel for schedule(dynamic, 4096)
    for (int64_t r = 0; r < (int64_t)n; ++r) { norm(out_hi[r]); norm(in_hi[r]); }
    uint64_t m_orig = 0;
    for (uint32_t r = 0; r < n; ++r) m_orig += out_hi[r].size() + in_hi[r].size();

    // Every fill arc joins two ancestors of the contracted node in the
    // elimination tree of the symmetrised original pattern (Liu), so a node's
    // lists are final once its subtree is contracted.  Phase A contracts the
    // small bottom subtrees in parallel (one thread per subtree, spin locks on
    // lists owned by nodes outside the subtree); phase B contracts the rest in
    // rank order with the pair loops spread over threads.
    std::vector<uint32_t> etp(n, INF);
    {
        std::vector<uint32_t> lo_off(n + 1, 0);
        for (uint32_t i = 0; i < n; ++i) {
            for (auto &e : out_hi[i]) lo_off[e.v + 1]++;
            for (auto &e : in_hi[i]) lo_off[e.v + 1]++;
        }
        for (uint32_t j = 0; j < n; ++j) lo_off[j + 1] += lo_off[j];
        std::vector<uint32_t> lo_adj(lo_off[n]);
        {
            std::vector<uint32_t> cur(lo_off.begin(), lo_off.end() - 1);
            for (uint32_t i = 0; i < n; ++i) {
                for (auto &e : out_hi[i]) lo_adj[cur[e.v]++] = i;
                for (auto &e : in_hi[i]) lo_adj[cur[e.v]++] = i;
            }
        }
        std::vector<uint32_t> anc(n, INF);
        for (uint32_t j = 0; j < n; ++j)
            for (uint32_t k = lo_off[j]; k < lo_off[j + 1]; ++k) {
                uint32_t r = lo_adj[k];
                while (anc[r] != INF && anc[r] != j) { uint32_t t = anc[r]; anc[r] = j; r = t; }
                if (anc[r] == INF) { anc[r] = j; etp[r] = j; }
            }
    }
    std::vector<uint32_t> sub(n, 1);
    for (uint32_t v = 0; v < n; ++v) if (etp[v] != INF) sub[etp[v]] += sub[v];
    const uint32_t SA = std::max<uint32_t>(4096, n / (uint32_t)(64 * std::max(1, omp_get_max_threads())));
    // A-root of every bottom node (INF for phase-B nodes); parents before children
    std::vector<uint32_t> aroot(n, INF);
    std::vector<uint32_t> roots;
    for (int64_t vv = (int64_t)n - 1; vv >= 0; --vv) {
        uint32_t v = (uint32_t)vv, p = etp[v];
        if (sub[v] > SA) continue;
        if (p == INF || sub[p] > SA) { aroot[v] = v; roots.push_back(v); }
        else aroot[v] = aroot[p];
    }
    // nodes of each A-subtree in ascending rank
    std::vector<uint32_t> rid(n, INF);
    for (size_t i = 0; i < roots.size(); ++i) rid[roots[i]] = (uint32_t)i;
    std::vector<uint32_t> ar_off(roots.size() + 1, 0), ar_nodes;
    for (uint32_t v = 0; v < n; ++v) if (aroot[v] != INF) ar_off[rid[aroot[v]] + 1]++;
    for (size_t i = 0; i < roots.size(); ++i) ar_off[i + 1] += ar_off[i];
    ar_nodes.resize(ar_off[roots.size()]);
    {
        std::vector<uint32_t> cur(ar_off.begin(), ar_off.end() - 1);
        for (uint32_t v = 0; v < n; ++v) if (aroot[v] != INF) ar_nodes[cur[rid[aroot[v]]]++] = v;
    }
    std::unique_ptr<std::atomic<uint8_t>[]> lk(new std::atomic<uint8_t>[n]);
    for (uint32_t v = 0; v < n; ++v) lk[v].store(0, std::memory_order_relaxed);
    auto lock = [&](uint32_t v) {
        while (lk[v].exchange(1, std::memory_order_acquire)) {
            while (lk[v].load(std::memory_order_relaxed)) { }
        }
    };
    auto unlock = [&](uint32_t v) { lk[v].store(0, std::memory_order_release); };
    auto contract_serial = [&](uint32_t x, uint32_t root, List &tmp) -> uint64_t {
        const List &I = in_hi[x];
        const List &O = out_hi[x];
        uint64_t fresh = 0;
        if (I.empty() || O.empty()) return 0;
        for (size_t a = 0; a < I.size(); ++a) {
            const uint32_t u = I[a].v, dux = I[a].w;
            auto it = std::upper_bound(O.begin(), O.end(), u, [](uint32_t val, const Ent &e) { return val < e.v; });
            size_t k0 = (size_t)(it - O.begin());
            if (k0 == O.size()) continue;
            const bool ext = aroot[u] != root;
            if (ext) lock(u);
            fresh += merge_add(out_hi[u], O.data() + k0, O.size() - k0, dux, tmp);
            if (ext) unlock(u);
        }
        for (size_t b = 0; b < O.size(); ++b) {
            const uint32_t w = O[b].v, dxw = O[b].w;
            auto it = std::upper_bound(I.begin(), I.end(), w, [](uint32_t val, const Ent &e) { return val < e.v; });
            size_t k0 = (size_t)(it - I.begin());
            if (k0 == I.size()) continue;
            const bool ext = aroot[w] != root;
            if (ext) lock(w);
            fresh += merge_add(in_hi[w], I.data() + k0, I.size() - k0, dxw, tmp);
            if (ext) unlock(w);
        }
        return fresh;
    };
    uint64_t n_short = 0;
    std::vector<List> tmps(omp_get_max_threads());
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : n_short)
    for (int64_t ri = 0; ri < (int64_t)roots.size(); ++ri) {
        List &tmp = tmps[omp_get_thread_num()];
        for (uint32_t k = ar_off[ri]; k < ar_off[ri + 1]; ++k) n_short += contract_serial(ar_nodes[k], roots[ri], tmp);
    }
    if (vb)
        fprintf(stderr, "[cchbuild] contract phase A: %zu subtrees (<= %u nodes), %zu nodes (%.1fs)\n", roots.size(),
                SA, ar_nodes.size(), now_s() - t2);
    size_t done_b = 0, n_b = n - ar_nodes.size(), report = std::max<size_t>(n_b / 10, 1);
    for (uint32_t x = 0; x < n; ++x) {
        if (aroot[x] != INF) continue;
        if (vb && ++done_b % report == 0)
            fprintf(stderr, "[cchbuild] contract phase B %5.1f%% (%.1fs)\n", 100.0 * done_b / n_b, now_s() - t2);
        const List &I = in_hi[x];
        const List &O = out_hi[x];
        if (I.empty() || O.empty()) continue;
        const size_t work = I.size() * O.size();
        if (work <= 65536) { n_short += contract_serial(x, INF - 1, tmps[0]); continue; }
        uint64_t fresh_total = 0;
        // (1) u in I: arcs u->w with w > u go to out_hi[u]
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : fresh_total)
        for (int64_t a = 0; a < (int64_t)I.size(); ++a) {
            const uint32_t u = I[a].v, dux = I[a].w;
            auto it = std::upper_bound(O.begin(), O.end(), u, [](uint32_t val, const Ent &e) { return val < e.v; });
            size_t k0 = (size_t)(it - O.begin());
            if (k0 == O.size()) continue;
            fresh_total += merge_add(out_hi[u], O.data() + k0, O.size() - k0, dux, tmps[omp_get_thread_num()]);
        }
        // (2) w in O: arcs u->w with u > w go to in_hi[w]
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : fresh_total)
        for (int64_t b = 0; b < (int64_t)O.size(); ++b) {
            const uint32_t w = O[b].v, dxw = O[b].w;
            auto it = std::upper_bound(I.begin(), I.end(), w, [](uint32_t val, const Ent &e) { return val < e.v; });
            size_t k0 = (size_t)(it - I.begin());
            if (k0 == I.size()) continue;
            fresh_total += merge_add(in_hi[w], I.data() + k0, I.size() - k0, dxw, tmps[omp_get_thread_num()]);
        }
        n_short += fresh_total;
    }
    double t3 = now_s();

    // ---- rank-aligned CSR ----
    bfc_cch *c = (bfc_cch *)calloc(1, sizeof(bfc_cch));
    c->n = n;
    c->up_off = (uint64_t *)malloc(sizeof(uint64_t) * (n + 1));
    c->dn_off = (uint64_t *)calloc(n + 1, sizeof(uint64_t));
    c->up_off[0] = 0;
    for (uint32_t r = 0; r < n; ++r) c->up_off[r + 1] = c->up_off[r] + out_hi[r].size();
    for (uint32_t x = 0; x < n; ++x)
        for (auto &e : in_hi[x]) c->dn_off[e.v + 1]++;
    for (uint32_t r = 0; r < n; ++r) c->dn_off[r + 1] += c->dn_off[r];
    c->m_up = c->up_off[n];
    c->m_down = c->dn_off[n];
    c->up_tgt = (uint32_t *)malloc(sizeof(uint32_t) * std::max<uint64_t>(c->m_up, 1));
    c->up_w = (uint32_t *)malloc(sizeof(uint32_t) * std::max<uint64_t>(c->m_up, 1));
    c->dn_tgt = (uint32_t *)malloc(sizeof(uint32_t) * std::max<uint64_t>(c->m_down, 1));
    c->dn_w = (uint32_t *)malloc(sizeof(uint32_t) * std::max<uint64_t>(c->m_down, 1));
#pragma omp parallel for schedule(dynamic, 4096)
    for (int64_t r = 0; r < (int64_t)n; ++r) {
        uint64_t o = c->up_off[r];
        for (auto &e : out_hi[r]) { c->up_tgt[o] = e.v; c->up_w[o] = e.w; ++o; }
    }
    {
        std::vector<uint64_t> cur(c->dn_off, c->dn_off + n);
        for (uint32_t x = 0; x < n; ++x)  // ascending x => per-tail targets ascending
            for (auto &e : in_hi[x]) { uint64_t o = cur[e.v]++; c->dn_tgt[o] = x; c->dn_w[o] = e.w; }
    }
    c->x = (float *)malloc(sizeof(float) * n);
    c->y = (float *)malloc(sizeof(float) * n);
    c->rank_to_filtered = (uint32_t *)malloc(sizeof(uint32_t) * n);
    for (uint32_t r = 0; r < n; ++r) {
        c->x[r] = g.x[inv_perm[r]];
        c->y[r] = g.y[inv_perm[r]];
        c->rank_to_filtered[r] = inv_perm[r];
    }
    // EBG in rank space
    c->m_ebg = g.head.size();
    c->ebg_off = (uint64_t *)malloc(sizeof(uint64_t) * (n + 1));
    c->ebg_head = (uint32_t *)malloc(sizeof(uint32_t) * std::max<uint64_t>(c->m_ebg, 1));
    c->ebg_w = (uint32_t *)malloc(sizeof(uint32_t) * std::max<uint64_t>(c->m_ebg, 1));
    c->ebg_off[0] = 0;
    for (uint32_t r = 0; r < n; ++r) {
        uint32_t f = inv_perm[r];
        c->ebg_off[r + 1] = c->ebg_off[r] + (g.off[f + 1] - g.off[f]);
    }
#pragma omp parallel for schedule(dynamic, 4096)
    for (int64_t r = 0; r < (int64_t)n; ++r) {
        uint32_t f = inv_perm[r];
        uint64_t o = c->ebg_off[r];
        for (uint64_t i = g.off[f]; i < g.off[f + 1]; ++i) { c->ebg_head[o] = perm[g.head[i]]; c->ebg_w[o] = g.w[i]; ++o; }
    }
    c->nbg_nodes = g.nbg_nodes;
    c->nbg_edges = g.nbg_edges;
    c->n_components = (uint32_t)comps.size();
    c->n_shortcuts = n_short;
    c->t_graph_s = t1 - t0;
    c->t_order_s = t2 - t1;
    c->t_contract_s = t3 - t2;
    if (vb) fprintf(stderr, "[cchbuild] contract+customize: orig %llu arcs, %llu shortcuts, up %llu down %llu (%.2fs)\n",
                    (unsigned long long)m_orig, (unsigned long long)n_short, (unsigned long long)c->m_up,
                    (unsigned long long)c->m_down, t3 - t2);
    return c;
}

}  // extern "C"
