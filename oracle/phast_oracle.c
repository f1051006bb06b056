/*
 * phast_oracle.c — CPU restatement of butterfly-route's PHAST engines.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity oracle for the MI355X
 * K-lane PHAST path: tests, the smoke check and the benchmark's CPU baseline
 * leg call it as the checker (or as the timed CPU reference); the product path
 * never links it.  It follows the reference's Rust line by line in plain C:
 *
 *   orc_query_batch            batched_phast.rs:131-260  (AoS, K <= 8)
 *   orc_query_batch_soa        batched_phast.rs:470-609 / :800-917 (SoA)
 *   orc_query_batch_bounded    batched_phast.rs:271-445  (512-rank lane bitsets)
 *   orc_query_batch_block_gated batched_phast.rs:946-1218 (4096 blocks, quirks)
 *   orc_compute_matrix_flat    batched_phast.rs:667-706  (chunks of 8, sequential)
 *   orc_compute_matrix_flat_soa batched_phast.rs:756-797
 *   orc_compute_matrix_block_gated batched_phast.rs:1223-1265
 *   orc_compute_matrix_flat_mt one K-batch per thread (the rayon per-source
 *                              parallelism of bucket_ch.rs:6816-6824)
 *   orc_phast_query            range/phast.rs:155-263 (single source)
 *   orc_phast_query_reverse    range/phast.rs:879-1108 semantics (all-to-one)
 *   orc_build_up_flat / orc_build_down_flat / orc_build_down_reverse_flat
 *                              matrix/bucket_ch.rs:191-424 (INF arcs dropped)
 *   orc_phast_seeded           server/isochrone_handler.rs:338-491
 *                              (generation stamps, 4096-rank block gating)
 *   orc_phast_seeded_2ch       server/isochrone_handler.rs:500-625 (time and
 *                              length-along-time, lexicographic tie-break)
 *   orc_table_bucket           matrix/bucket_ch.rs:3436-3520 (bucket many-to-many:
 *                              forward UP sweeps into buckets, backward sweeps over
 *                              DownReverseAdjFlat, join at every settled node)
 *   orc_dijkstra               plain Dijkstra, the brute-force pin
 *   orc_triangle_relax         customization.rs:1083-1200 (triangle_relax_parallel)
 *                              and :1253-1356 (the (time, length) relaxation of
 *                              triangle_relax_lex_parallel)
 *   orc_elect_middles          customization.rs:1358-1437 (its middle election)
 *
 * Parity pins (tests/): the reference's hand vectors (range/phast.rs:1198-1554,
 * and the customization vectors of customization.rs:2265-2500),
 * the broom closed form (bucket_ch.rs:5638-5670), brute-force Dijkstra on the
 * CCH union graph and on the edge-based graph, Floyd-Warshall in rank order for
 * the customized weights, and the invariants the
 * reference's own bench/validators assert (K-lane == single source, SoA == AoS,
 * bounded == unbounded within T, gated == ungated within T).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

#define INF 0xFFFFFFFFu
#define K_LANES 8

static inline uint32_t sat_add(uint32_t a, uint32_t b) {
    uint32_t s = a + b;
    return s < a ? INF : s;
}

/* ---- binary min-heap of (dist << 32 | node) ---- */
typedef struct { uint64_t *a; size_t n, cap; } heap_t;
static void hp_push(heap_t *h, uint32_t d, uint32_t v) {
    if (h->n == h->cap) {
        h->cap = h->cap ? h->cap * 2 : 256;
        h->a = (uint64_t *)realloc(h->a, h->cap * sizeof(uint64_t));
    }
    uint64_t x = ((uint64_t)d << 32) | v;
    size_t i = h->n++;
    while (i) {
        size_t p = (i - 1) / 2;
        if (h->a[p] <= x) break;
        h->a[i] = h->a[p];
        i = p;
    }
    h->a[i] = x;
}
static int hp_pop(heap_t *h, uint32_t *d, uint32_t *v) {
    if (!h->n) return 0;
    uint64_t top = h->a[0], x = h->a[--h->n];
    size_t i = 0;
    for (;;) {
        size_t l = 2 * i + 1, r = l + 1, m = i;
        uint64_t mv = x;
        if (l < h->n && h->a[l] < mv) { m = l; mv = h->a[l]; }
        if (r < h->n && h->a[r] < mv) { m = r; }
        if (m == i) break;
        h->a[i] = h->a[m];
        i = m;
    }
    if (h->n) h->a[i] = x;
    *d = (uint32_t)(top >> 32);
    *v = (uint32_t)top;
    return 1;
}
static void hp_free(heap_t *h) { free(h->a); h->a = NULL; h->n = h->cap = 0; }

typedef struct {
    uint32_t n;
    const uint64_t *up_off; const uint32_t *up_tgt; const uint32_t *up_w;
    const uint64_t *dn_off; const uint32_t *dn_tgt; const uint32_t *dn_w;
} orc_cch;

/* upward heap-Dijkstra of one lane into dist (stride `stride`, offset `lane`),
 * batched_phast.rs:156-189 */
static void up_dijkstra(const orc_cch *g, uint32_t origin, uint32_t *dist, size_t stride, size_t lane) {
    heap_t h = {0};
    hp_push(&h, 0, origin);
    uint32_t d, u;
    while (hp_pop(&h, &d, &u)) {
        if (d > dist[(size_t)u * stride + lane]) continue;
        for (uint64_t i = g->up_off[u]; i < g->up_off[u + 1]; ++i) {
            uint32_t v = g->up_tgt[i], w = g->up_w[i];
            if (w == INF) continue;
            uint32_t nd = sat_add(d, w);
            size_t vi = (size_t)v * stride + lane;
            if (nd < dist[vi]) { dist[vi] = nd; hp_push(&h, nd, v); }
        }
    }
    hp_free(&h);
}

/* batched_phast.rs:131-260 — AoS dist[lane*n + node], k <= 8 */
int orc_query_batch(const orc_cch *g, const uint32_t *src, uint32_t k, uint32_t *dist) {
    if (k > K_LANES) return 1;
    const size_t n = g->n;
    for (size_t i = 0; i < (size_t)k * n; ++i) dist[i] = INF;
    for (uint32_t l = 0; l < k; ++l) dist[l * n + src[l]] = 0;
    for (uint32_t l = 0; l < k; ++l) up_dijkstra(g, src[l], dist + l * n, 1, 0);
    for (int64_t r = (int64_t)n - 1; r >= 0; --r) {
        uint64_t a = g->dn_off[r], b = g->dn_off[r + 1];
        if (a == b) continue;
        int any = 0;
        for (uint32_t l = 0; l < k; ++l) if (dist[l * n + r] != INF) { any = 1; break; }
        if (!any) continue;
        for (uint64_t i = a; i < b; ++i) {
            uint32_t v = g->dn_tgt[i], w = g->dn_w[i];
            if (w == INF) continue;
            for (uint32_t l = 0; l < k; ++l) {
                uint32_t du = dist[l * n + r];
                if (du != INF) {
                    uint32_t nd = sat_add(du, w);
                    if (nd < dist[l * n + v]) dist[l * n + v] = nd;
                }
            }
        }
    }
    return 0;
}

/* batched_phast.rs:800-917 — SoA dist_soa[node*8 + lane], k <= 8 */
int orc_query_batch_soa(const orc_cch *g, const uint32_t *src, uint32_t k, uint32_t *dist_soa) {
    if (k > K_LANES) return 1;
    const size_t n = g->n;
    for (size_t i = 0; i < n * K_LANES; ++i) dist_soa[i] = INF;
    for (uint32_t l = 0; l < k; ++l) dist_soa[(size_t)src[l] * K_LANES + l] = 0;
    for (uint32_t l = 0; l < k; ++l) up_dijkstra(g, src[l], dist_soa, K_LANES, l);
    for (int64_t r = (int64_t)n - 1; r >= 0; --r) {
        uint64_t a = g->dn_off[r], b = g->dn_off[r + 1];
        if (a == b) continue;
        const uint32_t *du = dist_soa + (size_t)r * K_LANES;
        int any = 0;
        for (uint32_t l = 0; l < k; ++l) if (du[l] != INF) { any = 1; break; }
        if (!any) continue;
        uint32_t dl[K_LANES];
        for (uint32_t l = 0; l < K_LANES; ++l) dl[l] = l < k ? du[l] : INF;
        for (uint64_t i = a; i < b; ++i) {
            uint32_t v = g->dn_tgt[i], w = g->dn_w[i];
            if (w == INF) continue;
            uint32_t *dv = dist_soa + (size_t)v * K_LANES;
            for (uint32_t l = 0; l < k; ++l) {
                if (dl[l] != INF) {
                    uint32_t nd = sat_add(dl[l], w);
                    if (nd < dv[l]) dv[l] = nd;
                }
            }
        }
    }
    return 0;
}

/* batched_phast.rs:271-445 — bounded, per-lane 512-rank block bitsets.
 * Values <= threshold are exact; values above it are whatever the reference
 * leaves there (callers mask). */
int orc_query_batch_bounded(const orc_cch *g, const uint32_t *src, uint32_t k, uint32_t threshold, uint32_t *dist) {
    if (k > K_LANES) return 1;
    const size_t n = g->n;
    const size_t BS = 512, nb = (n + BS - 1) / BS, nw = (nb + 63) / 64;
    for (size_t i = 0; i < (size_t)k * n; ++i) dist[i] = INF;
    for (uint32_t l = 0; l < k; ++l) dist[l * n + src[l]] = 0;
    int lane_active[K_LANES] = {0};
    for (uint32_t l = 0; l < k; ++l) lane_active[l] = 1;
    uint64_t *ab = (uint64_t *)calloc((size_t)k * nw + 1, sizeof(uint64_t));
    for (uint32_t l = 0; l < k; ++l) { size_t b = src[l] / BS; ab[l * nw + b / 64] |= 1ull << (b % 64); }
    for (uint32_t l = 0; l < k; ++l) {
        heap_t h = {0};
        hp_push(&h, 0, src[l]);
        uint32_t d, u;
        uint32_t *dl = dist + l * n;
        while (hp_pop(&h, &d, &u)) {
            if (d > threshold) break;
            if (d > dl[u]) continue;
            for (uint64_t i = g->up_off[u]; i < g->up_off[u + 1]; ++i) {
                uint32_t v = g->up_tgt[i], w = g->up_w[i];
                if (w == INF) continue;
                uint32_t nd = sat_add(d, w);
                if (nd < dl[v]) {
                    dl[v] = nd;
                    hp_push(&h, nd, v);
                    if (nd <= threshold) { size_t b = v / BS; ab[l * nw + b / 64] |= 1ull << (b % 64); }
                }
            }
        }
        hp_free(&h);
    }
    for (uint32_t l = 0; l < k; ++l) {
        size_t cnt = 0;
        for (size_t w = 0; w < nw; ++w) cnt += (size_t)__builtin_popcountll(ab[l * nw + w]);
        if (cnt == 0) lane_active[l] = 0;
    }
    for (int64_t r = (int64_t)n - 1; r >= 0; --r) {
        size_t blk = (size_t)r / BS, wi = blk / 64;
        uint64_t bit = 1ull << (blk % 64);
        uint64_t a = g->dn_off[r], b = g->dn_off[r + 1];
        if (a == b) continue;
        int any = 0;
        for (uint32_t l = 0; l < k; ++l) if (lane_active[l] && (ab[l * nw + wi] & bit)) { any = 1; break; }
        if (!any) continue;
        for (uint64_t i = a; i < b; ++i) {
            uint32_t v = g->dn_tgt[i], w = g->dn_w[i];
            if (w == INF) continue;
            for (uint32_t l = 0; l < k; ++l) {
                if (!lane_active[l] || !(ab[l * nw + wi] & bit)) continue;
                uint32_t du = dist[l * n + r];
                if (du != INF) {
                    uint32_t nd = sat_add(du, w);
                    if (nd <= threshold && nd < dist[l * n + v]) {
                        dist[l * n + v] = nd;
                        size_t vb = v / BS;
                        ab[l * nw + vb / 64] |= 1ull << (vb % 64);
                    }
                }
            }
        }
    }
    free(ab);
    return 0;
}

/* batched_phast.rs:946-1218 — SoA + 4096-rank combined-active bitset, with
 * the reference's fallbacks (threshold == MAX, n < 16384, threshold >=
 * 300000) and the 30 % plain/gated switch.  AoS output dist[lane*n + node]. */
int orc_query_batch_block_gated(const orc_cch *g, const uint32_t *src, uint32_t k, uint32_t threshold, uint32_t *dist) {
    if (k > K_LANES) return 1;
    const size_t n = g->n, BS = 4096;
    if (threshold == INF || n < BS * 4 || threshold >= 300000u) return orc_query_batch(g, src, k, dist);
    uint32_t *ds = (uint32_t *)malloc(n * K_LANES * sizeof(uint32_t));
    for (size_t i = 0; i < n * K_LANES; ++i) ds[i] = INF;
    for (uint32_t l = 0; l < k; ++l) ds[(size_t)src[l] * K_LANES + l] = 0;
    const size_t nb = (n + BS - 1) / BS, nw = (nb + 63) / 64;
    uint64_t *ab = (uint64_t *)calloc((size_t)k * nw + 1, sizeof(uint64_t));
    for (uint32_t l = 0; l < k; ++l) { size_t b = src[l] / BS; ab[l * nw + b / 64] |= 1ull << (b % 64); }
    for (uint32_t l = 0; l < k; ++l) {
        heap_t h = {0};
        hp_push(&h, 0, src[l]);
        uint32_t d, u;
        while (hp_pop(&h, &d, &u)) {
            if (d > ds[(size_t)u * K_LANES + l]) continue;
            for (uint64_t i = g->up_off[u]; i < g->up_off[u + 1]; ++i) {
                uint32_t v = g->up_tgt[i], w = g->up_w[i];
                if (w == INF) continue;
                uint32_t nd = sat_add(d, w);
                size_t vi = (size_t)v * K_LANES + l;
                if (nd < ds[vi]) {
                    ds[vi] = nd;
                    hp_push(&h, nd, v);
                    if (nd <= threshold) { size_t b = v / BS; ab[l * nw + b / 64] |= 1ull << (b % 64); }
                }
            }
        }
        hp_free(&h);
    }
    uint64_t *comb = (uint64_t *)calloc(nw + 1, sizeof(uint64_t));
    for (uint32_t l = 0; l < k; ++l) for (size_t w = 0; w < nw; ++w) comb[w] |= ab[l * nw + w];
    size_t cnt = 0;
    for (size_t w = 0; w < nw; ++w) cnt += (size_t)__builtin_popcountll(comb[w]);
    double ratio = (double)cnt / (double)nb;
    if (ratio > 0.30) {
        for (int64_t r = (int64_t)n - 1; r >= 0; --r) {
            uint64_t a = g->dn_off[r], b = g->dn_off[r + 1];
            if (a == b) continue;
            const uint32_t *du = ds + (size_t)r * K_LANES;
            int any = 0;
            for (uint32_t l = 0; l < k; ++l) if (du[l] != INF) { any = 1; break; }
            if (!any) continue;
            uint32_t dl[K_LANES];
            for (uint32_t l = 0; l < K_LANES; ++l) dl[l] = l < k ? du[l] : INF;
            for (uint64_t i = a; i < b; ++i) {
                uint32_t v = g->dn_tgt[i], w = g->dn_w[i];
                if (w == INF) continue;
                for (uint32_t l = 0; l < k; ++l)
                    if (dl[l] != INF) {
                        uint32_t nd = sat_add(dl[l], w);
                        if (nd < ds[(size_t)v * K_LANES + l]) ds[(size_t)v * K_LANES + l] = nd;
                    }
            }
        }
    } else {
        for (int64_t bi = (int64_t)nb - 1; bi >= 0; --bi) {
            if (!((comb[bi / 64] >> (bi % 64)) & 1)) continue;
            size_t r0 = (size_t)bi * BS, r1 = r0 + BS < n ? r0 + BS : n;
            for (int64_t r = (int64_t)r1 - 1; r >= (int64_t)r0; --r) {
                uint64_t a = g->dn_off[r], b = g->dn_off[r + 1];
                if (a == b) continue;
                uint32_t dl[K_LANES];
                unsigned mask = 0;
                for (uint32_t l = 0; l < K_LANES; ++l) dl[l] = INF;
                for (uint32_t l = 0; l < k; ++l) {
                    dl[l] = ds[(size_t)r * K_LANES + l];
                    if (dl[l] != INF && dl[l] <= threshold) mask |= 1u << l;
                }
                if (!mask) continue;
                for (uint64_t i = a; i < b; ++i) {
                    uint32_t v = g->dn_tgt[i], w = g->dn_w[i];
                    if (w == INF) continue;
                    for (uint32_t l = 0; l < k; ++l) {
                        if (!((mask >> l) & 1)) continue;
                        uint32_t nd = sat_add(dl[l], w);
                        uint32_t *dv = ds + (size_t)v * K_LANES + l;
                        if (nd < *dv) {
                            *dv = nd;
                            if (nd <= threshold) {
                                size_t vb = v / BS;
                                ab[l * nw + vb / 64] |= 1ull << (vb % 64);
                                comb[vb / 64] |= 1ull << (vb % 64);
                            }
                        }
                    }
                }
            }
        }
    }
    for (uint32_t l = 0; l < k; ++l)
        for (size_t v = 0; v < n; ++v) dist[(size_t)l * n + v] = ds[v * K_LANES + l];
    free(ds); free(ab); free(comb);
    return 0;
}

/* batched_phast.rs:667-706 — sources.chunks(8) sequentially, row-major out */
int orc_compute_matrix_flat(const orc_cch *g, const uint32_t *src, uint32_t S, const uint32_t *tgt, uint32_t T,
                            uint32_t *out) {
    const size_t n = g->n;
    uint32_t *dist = (uint32_t *)malloc(n * K_LANES * sizeof(uint32_t));
    if (!dist) return 2;
    for (uint32_t b0 = 0; b0 < S; b0 += K_LANES) {
        uint32_t k = S - b0 < K_LANES ? S - b0 : K_LANES;
        orc_query_batch(g, src + b0, k, dist);
        for (uint32_t l = 0; l < k; ++l)
            for (uint32_t t = 0; t < T; ++t) out[(size_t)(b0 + l) * T + t] = dist[(size_t)l * n + tgt[t]];
    }
    free(dist);
    return 0;
}

/* batched_phast.rs:756-797 */
int orc_compute_matrix_flat_soa(const orc_cch *g, const uint32_t *src, uint32_t S, const uint32_t *tgt, uint32_t T,
                                uint32_t *out) {
    const size_t n = g->n;
    uint32_t *ds = (uint32_t *)malloc(n * K_LANES * sizeof(uint32_t));
    if (!ds) return 2;
    for (uint32_t b0 = 0; b0 < S; b0 += K_LANES) {
        uint32_t k = S - b0 < K_LANES ? S - b0 : K_LANES;
        orc_query_batch_soa(g, src + b0, k, ds);
        for (uint32_t l = 0; l < k; ++l)
            for (uint32_t t = 0; t < T; ++t) out[(size_t)(b0 + l) * T + t] = ds[(size_t)tgt[t] * K_LANES + l];
    }
    free(ds);
    return 0;
}

/* batched_phast.rs:1223-1265 */
int orc_compute_matrix_block_gated(const orc_cch *g, const uint32_t *src, uint32_t S, const uint32_t *tgt, uint32_t T,
                                   uint32_t threshold, uint32_t *out) {
    const size_t n = g->n;
    uint32_t *dist = (uint32_t *)malloc(n * K_LANES * sizeof(uint32_t));
    if (!dist) return 2;
    for (uint32_t b0 = 0; b0 < S; b0 += K_LANES) {
        uint32_t k = S - b0 < K_LANES ? S - b0 : K_LANES;
        orc_query_batch_block_gated(g, src + b0, k, threshold, dist);
        for (uint32_t l = 0; l < k; ++l)
            for (uint32_t t = 0; t < T; ++t) out[(size_t)(b0 + l) * T + t] = dist[(size_t)l * n + tgt[t]];
    }
    free(dist);
    return 0;
}

/* compute_matrix_flat_soa with one K-batch per thread (OpenMP); the CPU
 * all-cores baseline.  n_threads <= 0 uses the OpenMP default. */
int orc_compute_matrix_flat_mt(const orc_cch *g, const uint32_t *src, uint32_t S, const uint32_t *tgt, uint32_t T,
                               uint32_t *out, int n_threads) {
    const size_t n = g->n;
    const int64_t nb = ((int64_t)S + K_LANES - 1) / K_LANES;
    int err = 0;
#pragma omp parallel num_threads(n_threads > 0 ? n_threads : omp_get_max_threads()) if (nb > 1)
    {
        uint32_t *ds = (uint32_t *)malloc(n * K_LANES * sizeof(uint32_t));
        if (!ds) {
#pragma omp atomic write
            err = 2;
        }
#pragma omp for schedule(dynamic, 1)
        for (int64_t bi = 0; bi < nb; ++bi) {
            if (!ds) continue;
            uint32_t b0 = (uint32_t)bi * K_LANES;
            uint32_t k = S - b0 < K_LANES ? S - b0 : K_LANES;
            orc_query_batch_soa(g, src + b0, k, ds);
            for (uint32_t l = 0; l < k; ++l)
                for (uint32_t t = 0; t < T; ++t) out[(size_t)(b0 + l) * T + t] = ds[(size_t)tgt[t] * K_LANES + l];
        }
        free(ds);
    }
    return err;
}

/* range/phast.rs:155-263 */
int orc_phast_query(const orc_cch *g, uint32_t origin, uint32_t *dist) {
    const size_t n = g->n;
    for (size_t i = 0; i < n; ++i) dist[i] = INF;
    dist[origin] = 0;
    up_dijkstra(g, origin, dist, 1, 0);
    for (int64_t r = (int64_t)n - 1; r >= 0; --r) {
        uint32_t du = dist[r];
        if (du == INF) continue;
        for (uint64_t i = g->dn_off[r]; i < g->dn_off[r + 1]; ++i) {
            uint32_t v = g->dn_tgt[i], w = g->dn_w[i];
            if (w == INF) continue;
            uint32_t nd = sat_add(du, w);
            if (nd < dist[v]) dist[v] = nd;
        }
    }
    return 0;
}

/* ---- INF-filtered flats, bucket_ch.rs:191-424 ---- */
typedef struct { uint32_t n; uint64_t *off; uint32_t *tgt; uint32_t *w; } orc_flat;

static void flat_alloc(orc_flat *f, uint32_t n, uint64_t m) {
    f->n = n;
    f->off = (uint64_t *)calloc((size_t)n + 1, sizeof(uint64_t));
    f->tgt = (uint32_t *)malloc((m ? m : 1) * sizeof(uint32_t));
    f->w = (uint32_t *)malloc((m ? m : 1) * sizeof(uint32_t));
}

void orc_flat_free(orc_flat *f) { free(f->off); free(f->tgt); free(f->w); f->off = NULL; f->tgt = f->w = NULL; }

/* UpAdjFlat::build_with / DownAdjFlat::build — keyed by tail */
static int build_tail_flat(uint32_t n, const uint64_t *off, const uint32_t *tgt, const uint32_t *w, orc_flat *f) {
    uint64_t m = 0;
    for (uint64_t i = 0; i < off[n]; ++i) m += w[i] != INF;
    flat_alloc(f, n, m);
    uint64_t k = 0;
    for (uint32_t r = 0; r < n; ++r) {
        f->off[r] = k;
        for (uint64_t i = off[r]; i < off[r + 1]; ++i)
            if (w[i] != INF) { f->tgt[k] = tgt[i]; f->w[k] = w[i]; ++k; }
    }
    f->off[n] = k;
    return 0;
}
int orc_build_up_flat(const orc_cch *g, orc_flat *f) { return build_tail_flat(g->n, g->up_off, g->up_tgt, g->up_w, f); }
int orc_build_down_flat(const orc_cch *g, orc_flat *f) { return build_tail_flat(g->n, g->dn_off, g->dn_tgt, g->dn_w, f); }

/* DownReverseAdjFlat::build_with — keyed by head, sources in tail order */
int orc_build_down_reverse_flat(const orc_cch *g, orc_flat *f) {
    const uint32_t n = g->n;
    uint64_t m = 0;
    uint64_t *cnt = (uint64_t *)calloc((size_t)n + 1, sizeof(uint64_t));
    for (uint32_t r = 0; r < n; ++r)
        for (uint64_t i = g->dn_off[r]; i < g->dn_off[r + 1]; ++i)
            if (g->dn_w[i] != INF) { cnt[g->dn_tgt[i]]++; ++m; }
    flat_alloc(f, n, m);
    uint64_t run = 0;
    for (uint32_t v = 0; v < n; ++v) { f->off[v] = run; run += cnt[v]; cnt[v] = f->off[v]; }
    f->off[n] = run;
    for (uint32_t r = 0; r < n; ++r)
        for (uint64_t i = g->dn_off[r]; i < g->dn_off[r + 1]; ++i) {
            if (g->dn_w[i] == INF) continue;
            uint64_t p = cnt[g->dn_tgt[i]]++;
            f->tgt[p] = r;
            f->w[p] = g->dn_w[i];
        }
    free(cnt);
    return 0;
}

/* Reverse PHAST to `target` (all-to-one), range/phast.rs:879-1108: upward
 * Dijkstra over DOWN arcs reversed, then a rank-descending pull over UP arcs. */
int orc_phast_query_reverse(const orc_cch *g, uint32_t target, uint32_t *dist) {
    const uint32_t n = g->n;
    orc_flat rev;
    orc_build_down_reverse_flat(g, &rev);
    for (uint32_t i = 0; i < n; ++i) dist[i] = INF;
    dist[target] = 0;
    heap_t h = {0};
    hp_push(&h, 0, target);
    uint32_t d, u;
    while (hp_pop(&h, &d, &u)) {
        if (d > dist[u]) continue;
        for (uint64_t i = rev.off[u]; i < rev.off[u + 1]; ++i) {
            uint32_t v = rev.tgt[i];
            uint32_t nd = sat_add(d, rev.w[i]);
            if (nd < dist[v]) { dist[v] = nd; hp_push(&h, nd, v); }
        }
    }
    hp_free(&h);
    for (int64_t r = (int64_t)n - 1; r >= 0; --r) {
        for (uint64_t i = g->up_off[r]; i < g->up_off[r + 1]; ++i) {
            uint32_t w = g->up_w[i];
            if (w == INF) continue;
            uint32_t nd = sat_add(dist[g->up_tgt[i]], w);
            if (nd < dist[r]) dist[r] = nd;
        }
    }
    orc_flat_free(&rev);
    return 0;
}

/* server/isochrone_handler.rs:338-491 — seeded bounded PHAST over flats.
 * State (generation-stamped labels and 4096-rank block stamps) is owned by the
 * caller through orc_seeded_state so repeated calls reuse it, as the
 * reference's thread-local PhastState does.  Output: (rank, dist) pairs with
 * dist <= threshold in ascending rank; returns the count. */
typedef struct {
    uint32_t n, n_blocks, gen;
    uint32_t *dist, *version, *block_active;
    uint32_t *len; /* length channel, allocated by the first 2-channel query */
} orc_seeded_state;

orc_seeded_state *orc_seeded_state_new(uint32_t n) {
    orc_seeded_state *s = (orc_seeded_state *)calloc(1, sizeof(*s));
    s->n = n;
    s->n_blocks = (n + 4095) / 4096;
    s->dist = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
    s->version = (uint32_t *)calloc(n ? n : 1, sizeof(uint32_t));
    s->block_active = (uint32_t *)calloc(s->n_blocks ? s->n_blocks : 1, sizeof(uint32_t));
    for (uint32_t i = 0; i < n; ++i) s->dist[i] = INF;
    return s;
}
void orc_seeded_state_free(orc_seeded_state *s) {
    if (!s) return;
    free(s->dist); free(s->version); free(s->block_active); free(s->len); free(s);
}
static inline uint32_t st_get(const orc_seeded_state *s, uint32_t v) { return s->version[v] == s->gen ? s->dist[v] : INF; }
static inline void st_set(orc_seeded_state *s, uint32_t v, uint32_t d) {
    s->dist[v] = d;
    s->version[v] = s->gen;
    s->block_active[v / 4096] = s->gen;
}

uint64_t orc_phast_seeded(orc_seeded_state *s, const orc_flat *up, const orc_flat *down, const uint32_t *seed_rank,
                          const uint32_t *seed_cost, uint32_t n_seeds, uint32_t threshold, uint32_t *out_rank,
                          uint32_t *out_dist) {
    const uint32_t n = s->n;
    s->gen += 1;
    if (s->gen == 0) {
        memset(s->version, 0, sizeof(uint32_t) * n);
        memset(s->block_active, 0, sizeof(uint32_t) * s->n_blocks);
        s->gen = 1;
    }
    for (uint32_t i = 0; i < n_seeds; ++i)
        if (seed_cost[i] < st_get(s, seed_rank[i])) st_set(s, seed_rank[i], seed_cost[i]);
    heap_t h = {0};
    for (uint32_t i = 0; i < n_seeds; ++i)
        if (st_get(s, seed_rank[i]) == seed_cost[i]) hp_push(&h, seed_cost[i], seed_rank[i]);
    uint32_t d, u;
    while (hp_pop(&h, &d, &u)) {
        if (d > threshold) break;
        if (d > st_get(s, u)) continue;
        for (uint64_t i = up->off[u]; i < up->off[u + 1]; ++i) {
            uint32_t v = up->tgt[i];
            uint32_t nd = sat_add(d, up->w[i]);
            if (nd < st_get(s, v)) { st_set(s, v, nd); hp_push(&h, nd, v); }
        }
    }
    hp_free(&h);
    for (int64_t b = (int64_t)s->n_blocks - 1; b >= 0; --b) {
        if (s->block_active[b] != s->gen) continue;
        uint32_t r0 = (uint32_t)b * 4096, r1 = r0 + 4096 < n ? r0 + 4096 : n;
        for (int64_t r = (int64_t)r1 - 1; r >= (int64_t)r0; --r) {
            uint32_t du = st_get(s, (uint32_t)r);
            if (du == INF || du > threshold) continue;
            for (uint64_t i = down->off[r]; i < down->off[r + 1]; ++i) {
                uint32_t v = down->tgt[i];
                uint32_t nd = sat_add(du, down->w[i]);
                if (nd < st_get(s, v)) st_set(s, v, nd);
            }
        }
    }
    uint64_t cnt = 0;
    for (uint32_t b = 0; b < s->n_blocks; ++b) {
        if (s->block_active[b] != s->gen) continue;
        uint32_t r0 = b * 4096, r1 = r0 + 4096 < n ? r0 + 4096 : n;
        for (uint32_t r = r0; r < r1; ++r)
            if (s->version[r] == s->gen && s->dist[r] <= threshold) {
                if (out_rank) out_rank[cnt] = r;
                if (out_dist) out_dist[cnt] = s->dist[r];
                ++cnt;
            }
    }
    return cnt;
}

/* server/isochrone_handler.rs:500-625 — 2-channel seeded bounded PHAST: a
 * length-along-time channel carried with the time field.  The *_len flats
 * share the time flats' topology (index i is the same arc).  Time is primary;
 * a strict time improvement takes the new length, an equal finite time keeps
 * the shorter length (the lazy lexicographic tie-break of :551-565 in the
 * upward phase, which re-pushes the node, and :591-602 in the downward scan).
 * Seeds apply in order while strictly faster (:522-526), so among seeds of one
 * rank the first of the fastest wins.  Output: (rank, time, len) with
 * time <= threshold in ascending rank; returns the count. */
static inline uint32_t st_len(const orc_seeded_state *s, uint32_t v) { return s->version[v] == s->gen ? s->len[v] : INF; }
static inline void st_set2(orc_seeded_state *s, uint32_t v, uint32_t d, uint32_t l) {
    st_set(s, v, d);
    s->len[v] = l;
}

uint64_t orc_phast_seeded_2ch(orc_seeded_state *s, const orc_flat *up, const orc_flat *down, const orc_flat *up_len,
                              const orc_flat *down_len, const uint32_t *seed_rank, const uint32_t *seed_time,
                              const uint32_t *seed_len, uint32_t n_seeds, uint32_t threshold, uint32_t *out_rank,
                              uint32_t *out_time, uint32_t *out_len) {
    const uint32_t n = s->n;
    s->gen += 1;
    if (s->gen == 0) {
        memset(s->version, 0, sizeof(uint32_t) * n);
        memset(s->block_active, 0, sizeof(uint32_t) * s->n_blocks);
        s->gen = 1;
    }
    if (!s->len) { /* ensure_len, :167-171 */
        s->len = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
        for (uint32_t i = 0; i < n; ++i) s->len[i] = INF;
    }
    for (uint32_t i = 0; i < n_seeds; ++i) /* :522-526 */
        if (seed_time[i] < st_get(s, seed_rank[i])) st_set2(s, seed_rank[i], seed_time[i], seed_len[i]);
    heap_t h = {0};
    for (uint32_t i = 0; i < n_seeds; ++i) /* :528-532 */
        if (st_get(s, seed_rank[i]) == seed_time[i]) hp_push(&h, seed_time[i], seed_rank[i]);
    uint32_t d, u;
    while (hp_pop(&h, &d, &u)) { /* :533-568 */
        if (d > threshold) break;
        if (d > st_get(s, u)) continue;
        const uint32_t lu = st_len(s, u);
        for (uint64_t i = up->off[u]; i < up->off[u + 1]; ++i) {
            const uint32_t v = up->tgt[i];
            const uint32_t nt = sat_add(d, up->w[i]), ct = st_get(s, v);
            if (nt < ct) {
                st_set2(s, v, nt, sat_add(lu, up_len->w[i]));
                hp_push(&h, nt, v);
            } else if (nt == ct && ct != INF) {
                const uint32_t nl = sat_add(lu, up_len->w[i]);
                if (nl < st_len(s, v)) {
                    st_set2(s, v, nt, nl);
                    hp_push(&h, nt, v);
                }
            }
        }
    }
    hp_free(&h);
    for (int64_t b = (int64_t)s->n_blocks - 1; b >= 0; --b) { /* :569-604 */
        if (s->block_active[b] != s->gen) continue;
        uint32_t r0 = (uint32_t)b * 4096, r1 = r0 + 4096 < n ? r0 + 4096 : n;
        for (int64_t r = (int64_t)r1 - 1; r >= (int64_t)r0; --r) {
            const uint32_t du = st_get(s, (uint32_t)r);
            if (du == INF || du > threshold) continue;
            const uint32_t lu = st_len(s, (uint32_t)r);
            for (uint64_t i = down->off[r]; i < down->off[r + 1]; ++i) {
                const uint32_t v = down->tgt[i];
                const uint32_t nt = sat_add(du, down->w[i]), ct = st_get(s, v);
                if (nt < ct) {
                    st_set2(s, v, nt, sat_add(lu, down_len->w[i]));
                } else if (nt == ct && ct != INF) {
                    const uint32_t nl = sat_add(lu, down_len->w[i]);
                    if (nl < st_len(s, v)) st_set2(s, v, nt, nl);
                }
            }
        }
    }
    uint64_t cnt = 0;
    for (uint32_t b = 0; b < s->n_blocks; ++b) { /* :605-622 */
        if (s->block_active[b] != s->gen) continue;
        uint32_t r0 = b * 4096, r1 = r0 + 4096 < n ? r0 + 4096 : n;
        for (uint32_t r = r0; r < r1; ++r)
            if (s->version[r] == s->gen && s->dist[r] <= threshold) {
                if (out_rank) out_rank[cnt] = r;
                if (out_time) out_time[cnt] = s->dist[r];
                if (out_len) out_len[cnt] = s->len[r];
                ++cnt;
            }
    }
    return cnt;
}

/* matrix/bucket_ch.rs:3436-3520 (table_bucket_parallel_bounded) in one
 * thread: per source a heap-Dijkstra over the UP flat (forward_fill_buckets_flat,
 * :3240-3272) pushing (node, source, dist) for every settled node; the items
 * grouped by node (PrefixSumBuckets); per target a heap-Dijkstra over the
 * DOWN-reverse flat (backward_join_prefix, :3357-3416) joining
 * cell = min(cell, bucket dist + d) at every settled node.  A pop above
 * `threshold` ends a sweep (SearchState::pop, :2435-2450), so cells above it
 * are whatever the sweeps leave (callers mask).  out: row-major S x T. */
typedef struct { uint32_t node, src, dist; } bucket_item;

static void bucket_sweep(const orc_flat *f, uint32_t origin, uint32_t threshold, uint32_t *dist, uint32_t *touched,
                         size_t *nt, bucket_item **items, size_t *ni, size_t *cap, uint32_t tag, int collect,
                         const uint64_t *boff, const bucket_item *bk, uint32_t *row, uint32_t T, uint32_t col) {
    heap_t h = {0};
    *nt = 0;
    dist[origin] = 0;
    touched[(*nt)++] = origin;
    hp_push(&h, 0, origin);
    uint32_t d, u;
    while (hp_pop(&h, &d, &u)) {
        if (d > dist[u]) continue;  /* a stale heap entry: u settled earlier */
        if (d > threshold) break;
        if (collect) {
            if (*ni == *cap) {
                *cap = *cap ? *cap * 2 : 1024;
                *items = (bucket_item *)realloc(*items, *cap * sizeof(bucket_item));
            }
            (*items)[(*ni)++] = (bucket_item){u, tag, d};
        } else {
            for (uint64_t k = boff[u]; k < boff[u + 1]; ++k) {
                uint32_t *cell = row + (size_t)bk[k].src * T + col;
                uint32_t tot = sat_add(bk[k].dist, d);
                if (tot < *cell) *cell = tot;
            }
        }
        for (uint64_t i = f->off[u]; i < f->off[u + 1]; ++i) {
            uint32_t v = f->tgt[i], nd = sat_add(d, f->w[i]);
            if (nd < dist[v]) {
                if (dist[v] == INF) touched[(*nt)++] = v;
                dist[v] = nd;
                hp_push(&h, nd, v);
            }
        }
    }
    for (size_t k = 0; k < *nt; ++k) dist[touched[k]] = INF;
    hp_free(&h);
}

int orc_table_bucket(const orc_flat *up, const orc_flat *down_rev, const uint32_t *src, uint32_t S,
                     const uint32_t *tgt, uint32_t T, uint32_t threshold, uint32_t *out) {
    const uint32_t n = up->n;
    for (size_t i = 0; i < (size_t)S * T; ++i) out[i] = INF;
    uint32_t *dist = (uint32_t *)malloc((size_t)n * sizeof(uint32_t));
    uint32_t *touched = (uint32_t *)malloc((size_t)n * sizeof(uint32_t));
    for (uint32_t v = 0; v < n; ++v) dist[v] = INF;
    bucket_item *items = NULL;
    size_t ni = 0, cap = 0, nt = 0;
    for (uint32_t s = 0; s < S; ++s)
        bucket_sweep(up, src[s], threshold, dist, touched, &nt, &items, &ni, &cap, s, 1, NULL, NULL, NULL, 0, 0);
    /* PrefixSumBuckets: items grouped by node, in push order within a node */
    uint64_t *boff = (uint64_t *)calloc((size_t)n + 1, sizeof(uint64_t));
    for (size_t k = 0; k < ni; ++k) boff[items[k].node + 1]++;
    for (uint32_t v = 0; v < n; ++v) boff[v + 1] += boff[v];
    bucket_item *bk = (bucket_item *)malloc((ni ? ni : 1) * sizeof(bucket_item));
    uint64_t *fill = (uint64_t *)malloc(((size_t)n + 1) * sizeof(uint64_t));
    memcpy(fill, boff, ((size_t)n + 1) * sizeof(uint64_t));
    for (size_t k = 0; k < ni; ++k) bk[fill[items[k].node]++] = items[k];
    for (uint32_t t = 0; t < T; ++t)
        bucket_sweep(down_rev, tgt[t], threshold, dist, touched, &nt, NULL, NULL, NULL, 0, 0, boff, bk, out, T, t);
    free(fill); free(bk); free(boff); free(items); free(touched); free(dist);
    return 0;
}

/* Plain Dijkstra on a CSR graph: the brute-force pin. */
int orc_dijkstra(uint32_t n, const uint64_t *off, const uint32_t *head, const uint32_t *w, uint32_t src, uint32_t *dist) {
    for (uint32_t i = 0; i < n; ++i) dist[i] = INF;
    dist[src] = 0;
    heap_t h = {0};
    hp_push(&h, 0, src);
    uint32_t d, u;
    while (hp_pop(&h, &d, &u)) {
        if (d > dist[u]) continue;
        for (uint64_t i = off[u]; i < off[u + 1]; ++i) {
            if (w[i] == INF) continue;
            uint32_t nd = sat_add(d, w[i]);
            if (nd < dist[head[i]]) { dist[head[i]] = nd; hp_push(&h, nd, head[i]); }
        }
    }
    hp_free(&h);
    return 0;
}

/* ---- CCH customization: triangle relaxation ---------------------------- */

/* find_edge_index (customization.rs:1507-1520): binary search of v among the
 * ascending targets of u; 1 and *idx when found. */
static int find_edge_index(uint32_t u, uint32_t v, const uint64_t *off, const uint32_t *tgt, uint64_t *idx) {
    uint64_t lo = off[u], hi = off[u + 1];
    while (lo < hi) {
        uint64_t mid = lo + (hi - lo) / 2;
        if (tgt[mid] < v) lo = mid + 1;
        else hi = mid;
    }
    if (lo < off[u + 1] && tgt[lo] == v) { *idx = lo; return 1; }
    return 0;
}

/* ReverseDownAdj (customization.rs:832-905): per apex m the DOWN arcs x -> m
 * as (x, arc index), x ascending.  Caller frees *roff, *rsrc, *ridx. */
static void reverse_down_adj(uint32_t n, const uint64_t *dn_off, const uint32_t *dn_tgt, uint64_t **roff,
                             uint32_t **rsrc, uint64_t **ridx) {
    uint64_t md = dn_off[n];
    uint64_t *off = (uint64_t *)calloc((size_t)n + 1, sizeof(uint64_t));
    for (uint64_t i = 0; i < md; ++i) off[dn_tgt[i] + 1]++;
    for (uint32_t m = 0; m < n; ++m) off[m + 1] += off[m];
    uint64_t *ins = (uint64_t *)calloc((size_t)n + 1, sizeof(uint64_t));
    uint32_t *src = (uint32_t *)malloc((md ? md : 1) * sizeof(uint32_t));
    uint64_t *idx = (uint64_t *)malloc((md ? md : 1) * sizeof(uint64_t));
    for (uint32_t u = 0; u < n; ++u)
        for (uint64_t i = dn_off[u]; i < dn_off[u + 1]; ++i) {
            uint32_t m = dn_tgt[i];
            uint64_t pos = off[m] + ins[m]++;
            src[pos] = u;
            idx[pos] = i;
        }
    free(ins);
    *roff = off; *rsrc = src; *ridx = idx;
}

/* triangle_relax_parallel (customization.rs:1083-1200) with lex = 0: words
 * are (weight << 32 | middle) and a candidate through apex m is
 * (w(x->m) + w(m->y) saturating) << 32 | m.  With lex = 1 it is
 * triangle_relax_lex_parallel's relaxation (:1253-1356): words are
 * (time << 32 | length) and both channels add separately, saturating.
 * Passes run over every apex until one changes nothing (the reference runs a
 * pass's apexes in parallel with relaxed fetch_min; a pass here takes them in
 * rank order, one of those interleavings, and the fixpoint is the same).
 * Returns the number of passes, or -1 when the time-only variant has not
 * converged after 100 passes (the reference panics there, :1180-1186).
 * *updates (nullable) receives the sum of the passes' update counts. */
int orc_triangle_relax(uint32_t n, const uint64_t *up_off, const uint32_t *up_tgt, const uint64_t *dn_off,
                       const uint32_t *dn_tgt, uint64_t *up_pk, uint64_t *dn_pk, int lex, uint64_t *updates) {
    uint64_t *roff; uint32_t *rsrc; uint64_t *ridx;
    reverse_down_adj(n, dn_off, dn_tgt, &roff, &rsrc, &ridx);
    uint64_t total = 0;
    int pass = 0;
    for (;;) {
        ++pass;
        uint64_t pu = 0;
        for (uint32_t m = 0; m < n; ++m) {
            for (uint64_t ir = roff[m]; ir < roff[m + 1]; ++ir) {
                uint32_t x = rsrc[ir];
                uint64_t p_xm = dn_pk[ridx[ir]];
                uint32_t w_xm = (uint32_t)(p_xm >> 32);
                if (w_xm == INF) continue;
                for (uint64_t i_my = up_off[m]; i_my < up_off[m + 1]; ++i_my) {
                    uint32_t y = up_tgt[i_my];
                    if (y == x) continue;
                    uint64_t p_my = up_pk[i_my];
                    uint32_t w_my = (uint32_t)(p_my >> 32);
                    if (w_my == INF) continue;
                    uint32_t low = lex ? sat_add((uint32_t)p_xm, (uint32_t)p_my) : m;
                    uint64_t cand = ((uint64_t)sat_add(w_xm, w_my) << 32) | low;
                    uint64_t idx;
                    uint64_t *word = NULL;
                    if (y > x) {
                        if (find_edge_index(x, y, up_off, up_tgt, &idx)) word = &up_pk[idx];
                    } else if (find_edge_index(x, y, dn_off, dn_tgt, &idx)) {
                        word = &dn_pk[idx];
                    }
                    if (word && cand < *word) { *word = cand; ++pu; }
                }
            }
        }
        total += pu;
        if (pu == 0) break;
        if (!lex && pass >= 100) { pass = -1; break; }
    }
    free(roff); free(rsrc); free(ridx);
    if (updates) *updates = total;
    return pass;
}

/* The middle election after triangle_relax_lex_parallel's relaxation
 * (customization.rs:1358-1437): every shortcut arc (bit i % 64 of word i / 64
 * of *_sc) takes the least apex whose halves' packed sum equals its converged
 * word; one with no such apex, and every original arc, keeps mid_in (nullable:
 * u32::MAX). */
int orc_elect_middles(uint32_t n, const uint64_t *up_off, const uint32_t *up_tgt, const uint64_t *dn_off,
                      const uint32_t *dn_tgt, const uint64_t *up_pk, const uint64_t *dn_pk, const uint64_t *up_sc,
                      const uint64_t *dn_sc, const uint32_t *up_mid_in, const uint32_t *dn_mid_in, uint32_t *up_mid,
                      uint32_t *dn_mid) {
    uint64_t mu = up_off[n], md = dn_off[n];
    uint64_t *roff; uint32_t *rsrc; uint64_t *ridx;
    reverse_down_adj(n, dn_off, dn_tgt, &roff, &rsrc, &ridx);
    for (uint64_t i = 0; i < mu; ++i) up_mid[i] = INF;
    for (uint64_t i = 0; i < md; ++i) dn_mid[i] = INF;
    for (uint32_t m = 0; m < n; ++m)
        for (uint64_t ir = roff[m]; ir < roff[m + 1]; ++ir) {
            uint32_t x = rsrc[ir];
            uint64_t p_xm = dn_pk[ridx[ir]];
            uint32_t t_xm = (uint32_t)(p_xm >> 32);
            if (t_xm == INF) continue;
            for (uint64_t i_my = up_off[m]; i_my < up_off[m + 1]; ++i_my) {
                uint32_t y = up_tgt[i_my];
                if (y == x) continue;
                uint64_t p_my = up_pk[i_my];
                uint32_t t_my = (uint32_t)(p_my >> 32);
                if (t_my == INF) continue;
                uint64_t cand = ((uint64_t)sat_add(t_xm, t_my) << 32) | sat_add((uint32_t)p_xm, (uint32_t)p_my);
                uint64_t idx;
                if (y > x) {
                    if (find_edge_index(x, y, up_off, up_tgt, &idx) && ((up_sc[idx >> 6] >> (idx & 63)) & 1) &&
                        cand == up_pk[idx] && m < up_mid[idx])
                        up_mid[idx] = m;
                } else if (find_edge_index(x, y, dn_off, dn_tgt, &idx) && ((dn_sc[idx >> 6] >> (idx & 63)) & 1) &&
                           cand == dn_pk[idx] && m < dn_mid[idx]) {
                    dn_mid[idx] = m;
                }
            }
        }
    for (uint64_t i = 0; i < mu; ++i)
        if (up_mid[i] == INF) up_mid[i] = up_mid_in ? up_mid_in[i] : INF;
    for (uint64_t i = 0; i < md; ++i)
        if (dn_mid[i] == INF) dn_mid[i] = dn_mid_in ? dn_mid_in[i] : INF;
    free(roff); free(rsrc); free(ridx);
    return 0;
}
