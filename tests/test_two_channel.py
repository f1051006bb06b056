"""Two-channel seeded PHAST: time and the length along the time-optimal path
(isochrone_handler.rs:500-625, the 2-channel table of bucket_ch.rs:6381-6520).

The oracle's restatement is pinned first, on the CPU:
  * the reference's own tie vectors (isochrone_handler.rs:1809-1876, committed
    as tests/golden/two_channel_vectors.json);
  * the broom with lengths = 2 x times (bucket_ch.rs:5801-5845): length is
    exactly twice the time;
  * brute force: a lexicographic Dijkstra over the two-layer graph (up layer,
    then down layer) of the CCH, a different algorithm from the PHAST walk;
  * the time channel equals the one-channel seeded PHAST, and lengths equal to
    the times give lengths equal to the times.
The GPU path is then checked bit-exact against the oracle (tests marked gpu).
"""
import ctypes as C
import heapq
import json
import os
import sys

import numpy as np
import pytest

from conftest import GOLDEN, broom_graph
from oracle.oracle import INF, Oracle


def say(*a):
    print("[gpu-test]", *a, flush=True)
    sys.stdout.flush()


def sat(a):
    return min(a, INF)


def tie_cases():
    with open(os.path.join(GOLDEN, "two_channel_vectors.json")) as f:
        return json.load(f)["cases"]


def tie_graph(case):
    from bfgpu import CchGraph
    g = CchGraph.from_arcs(case["n"], [(a, b, t) for a, b, t, _ in case["up"]],
                           [(a, b, t) for a, b, t, _ in case["down"]])
    # from_arcs sorts by (tail, head); lengths follow the same order
    up_len = [ln for _, _, _, ln in sorted(case["up"])]
    dn_len = [ln for _, _, _, ln in sorted(case["down"])]
    return g, up_len, dn_len


def lengths_for(g, seed):
    """Length weights with the time flats' INF pattern: finite where the time is."""
    rng = np.random.default_rng(seed)
    ul = rng.integers(0, 4000, g.m_up).astype(np.uint32)
    dl = rng.integers(0, 4000, g.m_down).astype(np.uint32)
    ul[g.up_w == INF] = INF
    dl[g.dn_w == INF] = INF
    return ul, dl


def tie_heavy(g, seed):
    """The same topology with times in {1, 2, 3}: equal-time paths everywhere."""
    from bfgpu import CchGraph
    rng = np.random.default_rng(seed)
    uw = rng.integers(1, 4, g.m_up).astype(np.uint32)
    dw = rng.integers(1, 4, g.m_down).astype(np.uint32)
    uw[g.up_w == INF] = INF
    dw[g.dn_w == INF] = INF
    return CchGraph(n=g.n, up_off=g.up_off, up_tgt=g.up_tgt, up_w=uw, dn_off=g.dn_off, dn_tgt=g.dn_tgt, dn_w=dw)


def first_fastest(seeds):
    """isochrone_handler.rs:522-526: per rank the first of the fastest seeds."""
    best = {}
    for r, t, ln in seeds:
        if t != INF and (r not in best or t < best[r][0]):
            best[r] = (t, ln)
    return best


def brute_2ch(g, ul, dl, seeds, thr):
    """Least (time, length) over the up-down paths from the seeds, by a
    lexicographic Dijkstra over two layers: layer 0 follows UP arcs, layer 1
    DOWN arcs, and a node passes from layer 0 to 1 at no cost."""
    n = g.n
    lab = {}
    pq = []
    for r, (t, ln) in first_fastest(seeds).items():
        lab[(0, r)] = (t, ln)
        heapq.heappush(pq, (t, ln, 0, r))
    done = set()
    while pq:
        t, ln, layer, u = heapq.heappop(pq)
        if (layer, u) in done or lab.get((layer, u)) != (t, ln):
            continue
        done.add((layer, u))
        nxt = []
        if layer == 0:
            nxt.append((1, u, 0, 0))
            for i in range(int(g.up_off[u]), int(g.up_off[u + 1])):
                if g.up_w[i] != INF:
                    nxt.append((0, int(g.up_tgt[i]), int(g.up_w[i]), int(ul[i])))
        else:
            for i in range(int(g.dn_off[u]), int(g.dn_off[u + 1])):
                if g.dn_w[i] != INF:
                    nxt.append((1, int(g.dn_tgt[i]), int(g.dn_w[i]), int(dl[i])))
        for lay, v, wt, wl in nxt:
            c = (sat(t + wt), sat(ln + wl))
            if c[0] == INF:
                continue
            if (lay, v) not in lab or c < lab[(lay, v)]:
                lab[(lay, v)] = c
                heapq.heappush(pq, (c[0], c[1], lay, v))
    out = {}
    for v in range(n):
        c = lab.get((1, v))
        if c is not None and c[0] <= thr:
            out[v] = c
    return out


# ---------------------------------------------------------------- CPU: oracle pins

@pytest.mark.parametrize("case", tie_cases(), ids=lambda c: c["name"])
def test_reference_tie_vectors(case):
    g, ul, dl = tie_graph(case)
    o = Oracle(g)
    o.attach_length(ul, dl)
    got = {r: (t, ln) for r, t, ln in o.seeded_2ch([tuple(s) for s in case["seeds"]], case["threshold"])}
    for r, (t, ln) in case["expect"].items():
        assert got[int(r)] == (t, ln)


@pytest.mark.parametrize("thr", [INF, 200, 55])
def test_broom_length_is_twice_the_time(thr):
    g, wu, wd = broom_graph(12)
    o = Oracle(g)
    o.attach_length(2 * g.up_w, 2 * g.dn_w)
    for s in range(12):
        got = o.seeded_2ch([(s, 0, 0)], thr)
        assert got, "the seed itself is settled"
        for r, t, ln in got:
            assert ln == 2 * t and t <= thr
        t1 = dict(o.seeded([(s, 0)], thr))
        assert {r: t for r, t, _ in got} == t1


@pytest.mark.parametrize("thr", [INF, 400])
def test_matches_lexicographic_bruteforce(grid_small, thr):
    for gi, g in enumerate([grid_small, tie_heavy(grid_small, 7)]):
        ul, dl = lengths_for(g, 11 + gi)
        o = Oracle(g)
        o.attach_length(ul, dl)
        rng = np.random.default_rng(3 + gi)
        for _ in range(4):
            k = int(rng.integers(1, 4))
            seeds = [(int(r), int(c), int(c) * 2) for r, c in zip(rng.integers(0, g.n, k), rng.integers(0, 30, k))]
            seeds.append((seeds[0][0], seeds[0][1], 0))  # same rank and time: the first seed keeps its length
            got = {r: (t, ln) for r, t, ln in o.seeded_2ch(seeds, thr)}
            assert got == brute_2ch(g, ul, dl, seeds, thr)


def test_time_channel_and_equal_lengths(grid_small):
    g = grid_small
    o = Oracle(g)
    o.attach_length(g.up_w, g.dn_w)
    for s in (0, g.n // 3, g.n - 1):
        got = o.seeded_2ch([(s, 5, 5)], 900)
        assert all(t == ln for _, t, ln in got)
        assert [(r, t) for r, t, _ in got] == o.seeded([(s, 5)], 900)


def test_length_flats_must_share_topology(grid_small):
    g = grid_small
    ul, dl = lengths_for(g, 1)
    ul[np.nonzero(g.up_w != INF)[0][0]] = INF  # a finite-time arc without a length
    with pytest.raises(ValueError):
        Oracle(g).attach_length(ul, dl)


def test_table_2ch_broom():
    """table_phast_lopsided_2ch on the broom (bucket_ch.rs:5880-5912): the
    length matrix is twice the time matrix within the bound."""
    g, wu, wd = broom_graph(12)
    o = Oracle(g)
    o.attach_length(2 * g.up_w, 2 * g.dn_w)
    seeds = [[(i, 0, 0)] for i in range(12)]
    for thr in (INF, 200, 55):
        tm, lm, conf = o.table_phast_lopsided_2ch(seeds, seeds, thr)
        t1, c1 = o.table_phast_lopsided([[(i, 0)] for i in range(12)], [[(i, 0)] for i in range(12)], thr)
        assert np.array_equal(tm, t1) and np.array_equal(conf, c1)
        fin = tm != INF
        assert np.array_equal(lm[fin], 2 * tm[fin]) and np.all(lm[~fin] == INF)


def test_table_2ch_first_fastest_target_seed_gives_the_length():
    """bucket_ch.rs:6462-6473: over a target's seeds in order, a strictly
    smaller time replaces the best, so among seeds of equal time the first
    one's length part applies (not the smallest length).  On the broom with
    lengths = 2 x times the field at leaf j is (t, 2t), so the cell length is
    2t - (first seed's length part), saturating."""
    g, wu, wd = broom_graph(12)
    o = Oracle(g)
    o.attach_length(2 * g.up_w, 2 * g.dn_w)
    src = [[(0, 0, 0)], [(5, 0, 0)]]
    tgt = [[(7, 0, 5), (7, 0, 1)], [(7, 0, 1), (7, 0, 5)], [(9, 3, 1 << 31), (9, 3, 0)]]
    tm, lm, conf = o.table_phast_lopsided_2ch(src, tgt, INF)
    for i, s in enumerate((0, 5)):
        t7 = int(wu[s] + wd[7])
        t9 = int(wu[s] + wd[9])
        assert tm[i * 3 + 0] == t7 and lm[i * 3 + 0] == 2 * t7 - 5
        assert tm[i * 3 + 1] == t7 and lm[i * 3 + 1] == 2 * t7 - 1
        assert tm[i * 3 + 2] == t9 - 3 and lm[i * 3 + 2] == 0  # 2t - 2^31 saturates at 0
    assert not conf.any()


# ---------------------------------------------------------------- GPU parity

@pytest.mark.gpu
@pytest.mark.parametrize("case", tie_cases(), ids=lambda c: c["name"])
def test_gpu_reference_tie_vectors(case):
    from bfgpu import BatchedPhastEngine
    say("2ch tie", case["name"])
    g, ul, dl = tie_graph(case)
    with BatchedPhastEngine(g) as e:
        e.attach_length(ul, dl)
        got = {r: (t, ln) for r, t, ln in e.run_phast_bounded_fast_seeded_2ch(
            [tuple(s) for s in case["seeds"]], case["threshold"])}
    for r, (t, ln) in case["expect"].items():
        assert got[int(r)] == (t, ln)


@pytest.mark.gpu
@pytest.mark.parametrize("ties", [False, True])
def test_gpu_2ch_matches_oracle(grid_config1, ties):
    from bfgpu import BatchedPhastEngine
    g = tie_heavy(grid_config1, 5) if ties else grid_config1
    ul, dl = lengths_for(g, 21)
    say("2ch grid", g.n, "ties" if ties else "times")
    o = Oracle(g)
    o.attach_length(ul, dl)
    rng = np.random.default_rng(9)
    lanes, thrs = [], []
    for i in range(70):  # two lane groups, the second partial
        k = int(rng.integers(1, 4))
        seeds = [(int(r), int(c), int(c) * 3) for r, c in zip(rng.integers(0, g.n, k), rng.integers(0, 60, k))]
        if i % 5 == 0:
            seeds.append((seeds[0][0], seeds[0][1], 0))  # duplicate rank, equal time
        if i % 7 == 0:
            seeds.append((seeds[0][0], INF, 0))  # an INF seed applies nowhere
        lanes.append(seeds)
        thrs.append(INF if i % 3 == 0 else int(rng.integers(20, 400 if ties else 3000)))
    with BatchedPhastEngine(g) as e:
        e.attach_length(ul, dl)
        t, ln = e.phast_seeded_2ch(lanes, thrs)
        t1 = e.phast_seeded([[(r, c) for r, c, _ in s] for s in lanes], thrs)
    assert np.array_equal(t, t1), "time channel differs from the one-channel PHAST"
    for i in range(len(lanes)):
        et, el = o.seeded_2ch_rows(lanes[i], thrs[i])
        assert np.array_equal(t[i], et), f"lane {i}: times"
        assert np.array_equal(ln[i], el), f"lane {i}: lengths"


@pytest.mark.gpu
@pytest.mark.parametrize("thr", [INF, 1500])
def test_gpu_table_2ch_matches_oracle(grid_config1, thr):
    from bfgpu import BatchedPhastEngine
    g = grid_config1
    ul, dl = lengths_for(g, 31)
    say("2ch table", g.n, thr)
    o = Oracle(g)
    o.attach_length(ul, dl)
    rng = np.random.default_rng(17)

    def seedset(k):
        return [(int(r), int(p), int(q)) for r, p, q in zip(rng.integers(0, g.n, k), rng.integers(0, 90, k),
                                                             rng.integers(0, 900, k))]
    src = [seedset(int(rng.integers(1, 3))) for _ in range(70)]
    tgt = [seedset(int(rng.integers(1, 3))) for _ in range(90)]
    src[3], tgt[5] = [], []  # empty seed sets: INF rows / columns, no conflicts
    tgt[7] = [(src[0][0][0], 3, 4)]  # a same-edge conflict with source 0
    tgt[8] = [(tgt[9][0][0], 0, 50), (tgt[9][0][0], 0, 10)]  # equal times: the first seed's length
    with BatchedPhastEngine(g) as e:
        e.attach_length(ul, dl)
        t, ln, conf = e.table_seeded_bounded_routed_2ch(src, tgt, thr)
        t1, c1 = e.table_seeded_bounded_routed([[s[:2] for s in v] for v in src], [[s[:2] for s in v] for v in tgt],
                                               thr)
    et, el, ec = o.table_phast_lopsided_2ch(src, tgt, thr)
    assert conf[7] == 1 and np.array_equal(conf, ec)
    assert np.array_equal(t, et)
    assert np.array_equal(ln, el)
    assert np.array_equal(t, t1) and np.array_equal(conf, c1)


@pytest.mark.gpu
def test_gpu_table_2ch_broom():
    from bfgpu import BatchedPhastEngine
    g, wu, wd = broom_graph(12)
    seeds = [[(i, 0, 0)] for i in range(12)]
    o = Oracle(g)
    o.attach_length(2 * g.up_w, 2 * g.dn_w)
    with BatchedPhastEngine(g) as e:
        e.attach_length(2 * g.up_w, 2 * g.dn_w)
        for thr in (INF, 200, 55):
            t, ln, conf = e.table_seeded_bounded_routed_2ch(seeds, seeds, thr)
            et, el, ec = o.table_phast_lopsided_2ch(seeds, seeds, thr)
            assert np.array_equal(t, et) and np.array_equal(ln, el) and np.array_equal(conf, ec)


@pytest.mark.gpu
def test_gpu_2ch_batches(grid_small, monkeypatch):
    """One 64-lane group per batch (BFG_T2_GROUPS=1): 150 lanes run as three
    batches, rows and table cells land at their lanes' offsets."""
    from bfgpu import BatchedPhastEngine
    monkeypatch.setenv("BFG_T2_GROUPS", "1")
    g = grid_small
    ul, dl = lengths_for(g, 41)
    say("2ch batches", g.n)
    o = Oracle(g)
    o.attach_length(ul, dl)
    rng = np.random.default_rng(43)
    lanes = [[(int(r), int(c), int(c))] for r, c in zip(rng.integers(0, g.n, 150), rng.integers(0, 40, 150))]
    thrs = [INF if i % 2 else 700 for i in range(150)]
    tgt = [[(int(r), 2, 5)] for r in rng.integers(0, g.n, 33)]
    with BatchedPhastEngine(g) as e:
        e.attach_length(ul, dl)
        t, ln = e.phast_seeded_2ch(lanes, thrs)
        tt, tl, tc = e.table_seeded_bounded_routed_2ch(lanes, tgt, 900)
    for i in (0, 63, 64, 127, 128, 149):
        et, el = o.seeded_2ch_rows(lanes[i], thrs[i])
        assert np.array_equal(t[i], et) and np.array_equal(ln[i], el), f"lane {i}"
    et, el, ec = o.table_phast_lopsided_2ch(lanes, tgt, 900)
    assert np.array_equal(tt, et) and np.array_equal(tl, el) and np.array_equal(tc, ec)


@pytest.mark.gpu
def test_gpu_2ch_needs_lengths(grid_small):
    from bfgpu import BatchedPhastEngine
    from bfgpu.engine import BfgError
    with BatchedPhastEngine(grid_small) as e:
        with pytest.raises(BfgError):
            e.phast_seeded_2ch([[(0, 0, 0)]], [INF])
        ul, dl = lengths_for(grid_small, 2)
        with pytest.raises(ValueError):  # one length per arc
            e.attach_length(ul[:-1], dl)
        other = grid_small.transpose()  # another CCH: per-node arc counts differ
        d = other.desc()
        rc = e._lib.bfg_ctx_attach_length(e._h, C.byref(d), other.up_w.ctypes.data_as(C.POINTER(C.c_uint32)),
                                          other.dn_w.ctypes.data_as(C.POINTER(C.c_uint32)))
        assert rc == 1, "a different graph must be refused"
        e.attach_length(ul, dl)
        t, ln = e.phast_seeded_2ch([[(0, 0, 0)], []], [INF, INF])
        assert t[0][0] == 0 and ln[0][0] == 0
        assert np.all(t[1] == INF) and np.all(ln[1] == INF)
