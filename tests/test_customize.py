"""CCH customization (triangle relaxation, route/src/customization.rs:1083-1440).

CPU: the oracle's restatement of triangle_relax_parallel and
triangle_relax_lex_parallel is pinned by the reference's own vectors
(topo_4node, customization.rs:2255-2500, in tests/golden/customization_vectors.json),
by Floyd-Warshall in rank order (the customized weight of an arc x -> y is the
shortest x -> y path of the edge-based graph whose inner nodes all rank below
both ends), by the CPU builder's fused contraction (an independent
implementation of the same fixpoint), and by the definition of the elected
middle.  GPU: bfg_triangle_relax and bfg_triangle_relax_lex equal the oracle
bit for bit (weights and middles) and the builder's weights, up to the
6.7M-node graph (tests/test_z_fullsize.py).
"""
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN

from bfgpu import CchGraph, GridSpec, build_grid, original_arc_weights
from bfgpu.customize import triangle_relax_lex_parallel, triangle_relax_parallel
from oracle import oracle as O

INF = 0xFFFFFFFF


def vectors():
    with open(os.path.join(GOLDEN, "customization_vectors.json")) as f:
        return json.load(f)


def topo4():
    t = vectors()["topo_4node"]
    g = CchGraph(n=t["n"], up_off=t["up_offsets"], up_tgt=t["up_targets"], up_w=np.zeros(3),
                 dn_off=t["down_offsets"], dn_tgt=t["down_targets"], dn_w=np.zeros(2))
    return g, t


_GRIDS = {}


def grid(rows, chord=0.0, seed=7):
    key = (rows, chord, seed)
    if key not in _GRIDS:
        _GRIDS[key] = build_grid(GridSpec(rows=rows, cols=rows, seed=seed, chord_frac=chord))
    return _GRIDS[key]


def floyd_rank_order(g):
    """Per CCH arc, the shortest path of the edge-based graph between its ends
    whose inner nodes all rank below both (Floyd-Warshall with intermediates
    taken in rank order, read for an arc before its lower end joins them)."""
    n = g.n
    D = np.full((n, n), INF, dtype=np.int64)
    tails = np.repeat(np.arange(n), np.diff(g.ebg_off.astype(np.int64)))
    np.minimum.at(D, (tails, g.ebg_head.astype(np.int64)), g.ebg_w.astype(np.int64))
    up_t = np.repeat(np.arange(n), np.diff(g.up_off.astype(np.int64)))
    dn_t = np.repeat(np.arange(n), np.diff(g.dn_off.astype(np.int64)))
    up_lo, dn_lo = up_t, g.dn_tgt.astype(np.int64)
    exp_up = np.empty(g.m_up, np.int64)
    exp_dn = np.empty(g.m_down, np.int64)
    by_up = [[] for _ in range(n)]
    by_dn = [[] for _ in range(n)]
    for a in range(g.m_up):
        by_up[up_lo[a]].append(a)
    for a in range(g.m_down):
        by_dn[dn_lo[a]].append(a)
    for k in range(n):
        for a in by_up[k]:
            exp_up[a] = D[up_t[a], g.up_tgt[a]]
        for a in by_dn[k]:
            exp_dn[a] = D[dn_t[a], g.dn_tgt[a]]
        D = np.minimum(D, np.minimum(D[:, k:k + 1] + D[k:k + 1, :], INF))
    return exp_up.astype(np.uint32), exp_dn.astype(np.uint32)


def middle_by_definition(g, up_w, dn_w, up_init_mid, dn_init_mid):
    """The (weight, middle) word of every arc as the least of its initial word
    and of (w(x -> m) + w(m -> y), m) over every apex m below both ends with
    both halves finite, found by listing each arc's apexes."""
    n = g.n
    up = {}
    for u in range(n):
        for i in range(int(g.up_off[u]), int(g.up_off[u + 1])):
            up[(u, int(g.up_tgt[i]))] = i
    dn = {}
    for u in range(n):
        for i in range(int(g.dn_off[u]), int(g.dn_off[u + 1])):
            dn[(u, int(g.dn_tgt[i]))] = i
    best_up = [(int(w) << 32) | int(m) for w, m in zip(up_w, up_init_mid)]
    best_dn = [(int(w) << 32) | int(m) for w, m in zip(dn_w, dn_init_mid)]
    into = [[] for _ in range(n)]  # apex m: DOWN arcs x -> m
    for (x, m), i in dn.items():
        into[m].append((x, i))
    for m in range(n):
        for x, ixm in into[m]:
            if dn_w[ixm] == INF:
                continue
            for j in range(int(g.up_off[m]), int(g.up_off[m + 1])):
                y = int(g.up_tgt[j])
                if y == x or up_w[j] == INF:
                    continue
                cand = (min(int(dn_w[ixm]) + int(up_w[j]), INF) << 32) | m
                if y > x and (x, y) in up:
                    best_up[up[(x, y)]] = min(best_up[up[(x, y)]], cand)
                elif y < x and (x, y) in dn:
                    best_dn[dn[(x, y)]] = min(best_dn[dn[(x, y)]], cand)
    return (np.array([b & 0xFFFFFFFF for b in best_up], np.uint32),
            np.array([b & 0xFFFFFFFF for b in best_dn], np.uint32))


# ---- the oracle, pinned (CPU) ------------------------------------------------

def test_oracle_lex_election_reference_vectors():
    g, t = topo4()
    for case in vectors()["lex_cases"]:
        for _ in range(2):  # deterministic: the same result twice
            tu, td, um, dm, passes, _, _ = O.triangle_relax_lex(
                g, case["time_up"], case["time_down"], case["len_up"], case["len_down"], t["up_is_shortcut"],
                t["down_is_shortcut"], t["up_middle"], t["down_middle"])
            assert tu[2] == case["expect_up_time_2"], case["name"]
            assert um[2] == case["expect_up_middle_2"], case["name"]
            # original arcs keep their weights and their given middles
            assert list(tu[:2]) == case["time_up"][:2] and list(td) == case["time_down"], case["name"]
            assert list(um[:2]) == [INF, INF] and list(dm) == [INF, INF], case["name"]


def test_oracle_time_relaxation_on_topo4():
    """The time-only relaxation on topo_4node: the shortcut 2 -> 3 through
    apex 0 (3 + 7) or apex 1 (5 + 5) both give 10, so apex 0 is elected; a
    seed below every triangle stays."""
    g, _ = topo4()
    uw, dw, um, dm, _ = O.triangle_relax(g, [7, 5, INF], [3, 5])
    assert list(uw) == [7, 5, 10] and list(um) == [INF, INF, 0]
    assert list(dw) == [3, 5] and list(dm) == [INF, INF]
    uw, _, um, _, _ = O.triangle_relax(g, [7, 5, 9], [3, 5], [INF, INF, 1], [INF, INF])
    assert uw[2] == 9 and um[2] == 1
    uw, _, um, _, _ = O.triangle_relax(g, [7, 4, INF], [3, 5])
    assert uw[2] == 9 and um[2] == 1


def test_pack_order_reference_vectors():
    for (a, b) in vectors()["pack_order"]["less_pairs"]:
        assert (a[0] << 32 | a[1]) < (b[0] << 32 | b[1])


@pytest.mark.parametrize("rows,chord,seed", [(10, 0.0, 7), (12, 0.0, 3), (14, 0.05, 11)])
def test_oracle_relax_equals_floyd_warshall_and_builder(rows, chord, seed):
    g = grid(rows, chord, seed)
    uw, dw, usc, dsc = original_arc_weights(g)
    ru, rd, mu, md, passes = O.triangle_relax(g, uw, dw)
    assert passes >= 1
    fu, fd = floyd_rank_order(g)
    assert np.array_equal(ru, fu) and np.array_equal(rd, fd)
    assert np.array_equal(ru, g.up_w) and np.array_equal(rd, g.dn_w)
    # every CCH arc the builder made that is not an edge-based arc is a shortcut
    assert usc.sum() + dsc.sum() > 0


@pytest.mark.parametrize("rows,seed", [(9, 5), (11, 2)])
def test_oracle_middles_by_definition(rows, seed):
    g = grid(rows, 0.0, seed)
    uw, dw, _, _ = original_arc_weights(g)
    ru, rd, mu, md, _ = O.triangle_relax(g, uw, dw)
    eu, ed = middle_by_definition(g, ru, rd, np.full(g.m_up, INF), np.full(g.m_down, INF))
    assert np.array_equal(mu, eu) and np.array_equal(md, ed)
    # a relaxed arc's middle splits it into two arcs whose weights add up to it
    for i in np.flatnonzero(mu != INF)[:200]:
        x = int(np.searchsorted(g.up_off, i, side="right") - 1)
        y, m = int(g.up_tgt[i]), int(mu[i])
        a = int(g.dn_off[x]) + int(np.searchsorted(g.dn_tgt[g.dn_off[x]:g.dn_off[x + 1]], m))
        b = int(g.up_off[m]) + int(np.searchsorted(g.up_tgt[g.up_off[m]:g.up_off[m + 1]], y))
        assert g.dn_tgt[a] == m and g.up_tgt[b] == y
        assert min(int(rd[a]) + int(ru[b]), INF) == ru[i]


def test_oracle_lex_election_matches_time_relaxation():
    """The (time, length) variant's times equal the time-only relaxation's
    (time is the primary key, customization.rs:1223-1227), and each elected
    middle of a shortcut gives the converged (time, length) as its two halves'
    sum."""
    g = grid(12, 0.0, 9)
    uw, dw, usc, dsc = original_arc_weights(g)
    rng = np.random.default_rng(4)
    lu = np.where(uw == INF, INF, rng.integers(1, 500, g.m_up)).astype(np.uint32)
    ld = np.where(dw == INF, INF, rng.integers(1, 500, g.m_down)).astype(np.uint32)
    tu, td, um, dm, _, cl_u, cl_d = O.triangle_relax_lex(g, uw, dw, lu, ld, usc, dsc)
    ru, rd, _, _, _ = O.triangle_relax(g, uw, dw)
    assert np.array_equal(tu, ru) and np.array_equal(td, rd)
    assert np.all(um[~usc] == INF) and np.all(dm[~dsc] == INF)
    for i in np.flatnonzero(usc & (um != INF))[:200]:
        x = int(np.searchsorted(g.up_off, i, side="right") - 1)
        y, m = int(g.up_tgt[i]), int(um[i])
        a = int(g.dn_off[x]) + int(np.searchsorted(g.dn_tgt[g.dn_off[x]:g.dn_off[x + 1]], m))
        b = int(g.up_off[m]) + int(np.searchsorted(g.up_tgt[g.up_off[m]:g.up_off[m + 1]], y))
        assert min(int(td[a]) + int(tu[b]), INF) == tu[i]
        assert min(int(cl_d[a]) + int(cl_u[b]), INF) == cl_u[i]


def test_original_arc_weights_of_the_builder_graph():
    g = grid(10)
    uw, dw, usc, dsc = original_arc_weights(g)
    # an original arc's weight is never below its customized weight, and relaxing
    # only ever lowers it
    assert np.all(g.up_w <= uw) and np.all(g.dn_w <= dw)
    assert np.all(uw[usc] == INF) and np.all(dw[dsc] == INF)


def test_relax_rejects_bad_inputs():
    g, _ = topo4()
    with pytest.raises(ValueError):
        triangle_relax_parallel(g, [1, 2], [3, 5])
    bad = CchGraph(n=4, up_off=[0, 1, 2, 3, 3], up_tgt=[3, 3, 1], up_w=np.zeros(3),
                   dn_off=[0, 0, 0, 2, 2], dn_tgt=[0, 1], dn_w=np.zeros(2))
    from bfgpu import BfgError
    with pytest.raises(BfgError) as e:  # an UP target below its tail, refused before any device call
        triangle_relax_parallel(bad, [1, 2, 3], [3, 5])
    assert e.value.code == 1
    # offsets that fall back: refused before any arc is read
    bad = CchGraph(n=4, up_off=[0, 3, 1, 3, 3], up_tgt=[3, 3, 3], up_w=np.zeros(3),
                   dn_off=[0, 0, 0, 2, 2], dn_tgt=[0, 1], dn_w=np.zeros(2))
    with pytest.raises(BfgError) as e:
        triangle_relax_parallel(bad, [1, 2, 3], [3, 5])
    assert e.value.code == 1 and "offsets" in str(e.value)


# ---- the GPU path --------------------------------------------------------------

@pytest.mark.gpu
def test_gpu_relax_reference_vectors():
    g, t = topo4()
    for case in vectors()["lex_cases"]:
        r = triangle_relax_lex_parallel(g, case["time_up"], case["time_down"], case["len_up"], case["len_down"],
                                        t["up_is_shortcut"], t["down_is_shortcut"], t["up_middle"],
                                        t["down_middle"])
        assert r.up_weights[2] == case["expect_up_time_2"], case["name"]
        assert r.up_middles[2] == case["expect_up_middle_2"], case["name"]
        assert list(r.up_middles[:2]) == [INF, INF] and list(r.down_middles) == [INF, INF]
        assert r.stats.passes == 1
    r = triangle_relax_parallel(g, [7, 5, INF], [3, 5])
    assert list(r.up_weights) == [7, 5, 10] and list(r.up_middles) == [INF, INF, 0]
    r = triangle_relax_parallel(g, [7, 5, 9], [3, 5], [INF, INF, 1], [INF, INF])
    assert r.up_weights[2] == 9 and r.up_middles[2] == 1


@pytest.mark.gpu
@pytest.mark.parametrize("rows,chord,seed", [(12, 0.0, 3), (32, 0.0, 20261203), (40, 0.05, 5)])
def test_gpu_relax_equals_oracle(rows, chord, seed):
    g = grid(rows, chord, seed)
    uw, dw, usc, dsc = original_arc_weights(g)
    print(f"[relax] grid {rows} n={g.n} m_up={g.m_up} m_down={g.m_down}", flush=True)
    r = triangle_relax_parallel(g, uw, dw)
    ru, rd, mu, md, _ = O.triangle_relax(g, uw, dw)
    assert np.array_equal(r.up_weights, ru) and np.array_equal(r.down_weights, rd)
    assert np.array_equal(r.up_middles, mu) and np.array_equal(r.down_middles, md)
    assert np.array_equal(r.up_weights, g.up_w) and np.array_equal(r.down_weights, g.dn_w)
    assert r.stats.triangles > 0 and r.stats.passes == 1
    # shortcuts seeded above their fixpoint with some middle (as the reference's
    # bottom-up weights are) relax to the same words as from INF
    rng = np.random.default_rng(seed)
    su = np.where(usc, ru.astype(np.uint64) + rng.integers(0, 50, g.m_up), uw).clip(0, INF).astype(np.uint32)
    sd = np.where(dsc, rd.astype(np.uint64) + rng.integers(0, 50, g.m_down), dw).clip(0, INF).astype(np.uint32)
    smu = np.where(usc, rng.integers(0, g.n, g.m_up), INF).astype(np.uint32)
    smd = np.where(dsc, rng.integers(0, g.n, g.m_down), INF).astype(np.uint32)
    r2 = triangle_relax_parallel(g, su, sd, smu, smd)
    e2 = O.triangle_relax(g, su, sd, smu, smd)
    assert np.array_equal(r2.up_weights, e2[0]) and np.array_equal(r2.down_weights, e2[1])
    assert np.array_equal(r2.up_middles, e2[2]) and np.array_equal(r2.down_middles, e2[3])
    assert np.array_equal(r2.up_weights, ru) and np.array_equal(r2.down_weights, rd)
    # lex: (time, length) with synthetic lengths
    lu = np.where(uw == INF, INF, rng.integers(1, 2000, g.m_up)).astype(np.uint32)
    ld = np.where(dw == INF, INF, rng.integers(1, 2000, g.m_down)).astype(np.uint32)
    rl = triangle_relax_lex_parallel(g, uw, dw, lu, ld, usc, dsc)
    el = O.triangle_relax_lex(g, uw, dw, lu, ld, usc, dsc)
    assert np.array_equal(rl.up_weights, el[0]) and np.array_equal(rl.down_weights, el[1])
    assert np.array_equal(rl.up_middles, el[2]) and np.array_equal(rl.down_middles, el[3])


@pytest.mark.gpu
def test_gpu_customize_grid_224_equals_builder():
    """The 224 x 224 grid (n = 166 k, 754 M triangles): the GPU's relaxation of
    the original weights equals the builder's customized weights."""
    from bfgpu import customize
    g = grid(224, 0.0, 20261203)
    print(f"[relax] grid 224 n={g.n} m_up={g.m_up} m_down={g.m_down}", flush=True)
    out, r = customize(g)
    print(f"[relax] {r.stats}", flush=True)
    assert np.array_equal(out.up_w, g.up_w) and np.array_equal(out.dn_w, g.dn_w)
    # every relaxed shortcut's middle splits it into two arcs that add up to it
    mu = r.up_middles
    rng = np.random.default_rng(1)
    pick = rng.choice(np.flatnonzero(mu != INF), 2000, replace=False)
    for i in pick:
        x = int(np.searchsorted(g.up_off, i, side="right") - 1)
        y, m = int(g.up_tgt[i]), int(mu[i])
        a = int(g.dn_off[x]) + int(np.searchsorted(g.dn_tgt[g.dn_off[x]:g.dn_off[x + 1]], m))
        b = int(g.up_off[m]) + int(np.searchsorted(g.up_tgt[g.up_off[m]:g.up_off[m + 1]], y))
        assert g.dn_tgt[a] == m and g.up_tgt[b] == y
        assert min(int(g.dn_w[a]) + int(g.up_w[b]), INF) == g.up_w[i]
