"""GPU parity: the HIP path through the C ABI against the CPU oracle.

Bit-exact everywhere (u32 integer work).  Sizes are chosen so the oracle
finishes in seconds; large shapes use size-independent properties (broom
closed form)."""
import sys

import numpy as np
import pytest

from conftest import broom_expect, broom_graph, grid_graph, hand_graph
from oracle.oracle import INF, Oracle

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-test]", *a, flush=True)
    sys.stdout.flush()


def mask(d, thr):
    d = np.array(d, dtype=np.uint32, copy=True)
    d[d > thr] = INF
    return d


@pytest.fixture(scope="module")
def engine_cls():
    from bfgpu import BatchedPhastEngine
    return BatchedPhastEngine


@pytest.mark.parametrize("name", ["five_node", "asymmetric_three_node", "inf_arc_four_node"])
def test_hand_vectors(engine_cls, name):
    say("hand", name)
    g, hv = hand_graph(name)
    o = Oracle(g)
    with engine_cls(g) as e:
        allsrc = list(range(g.n))
        r = e.query_batch(allsrc)
        for s in allsrc:
            assert np.array_equal(r.dist[s], o.phast_query(s))
        for s, exp in hv.get("forward", {}).items():
            assert r.dist[int(s)].tolist() == exp
        m, _ = e.compute_matrix_flat(allsrc, allsrc)
        assert np.array_equal(m, o.compute_matrix_flat(allsrc, allsrc))


@pytest.mark.parametrize("n_leaf", [12, 200_000])
def test_broom_closed_form(engine_cls, n_leaf):
    say("broom", n_leaf)
    g, wu, wd = broom_graph(n_leaf)
    rng = np.random.default_rng(n_leaf)
    s = rng.choice(n_leaf, min(n_leaf, 300), replace=False)
    t = rng.choice(n_leaf, min(n_leaf, 500), replace=False)
    t[0] = s[0]  # a diagonal cell
    with engine_cls(g) as e:
        m, _ = e.compute_matrix_flat(s, t)
        assert np.array_equal(m.reshape(s.size, t.size), broom_expect(wu, wd, s, t))
        for thr in (0, 75, 130, 100_000):
            mb, _ = e.compute_matrix_block_gated(s, t, thr)
            assert np.array_equal(mb.reshape(s.size, t.size), mask(broom_expect(wu, wd, s, t), thr))


def test_matrix_config1_100x100(engine_cls, grid_config1):
    """224x224 grid (the reference's CPU configuration), 100x100 uniform ranks."""
    g = grid_config1
    say("config1 100x100 n", g.n)
    rng = np.random.default_rng(20261203)
    s = rng.integers(0, g.n, 100).astype(np.uint32)
    t = rng.integers(0, g.n, 100).astype(np.uint32)
    o = Oracle(g)
    exp = o.compute_matrix_flat(s, t)
    with engine_cls(g) as e:
        got, st = e.compute_matrix_flat(s, t)
    assert np.array_equal(got, exp)
    assert (got != INF).mean() > 0.5


def root_path(g, r):
    """Ranks from r to its elimination-tree root: in a chordal CCH the parent of
    a node is its lowest-ranked UP neighbour."""
    out = [int(r)]
    while True:
        a, b = int(g.up_off[out[-1]]), int(g.up_off[out[-1] + 1])
        fin = g.up_tgt[a:b][g.up_w[a:b] != INF]
        if len(fin) == 0:
            return out
        out.append(int(fin.min()))


def test_matrix_target_paths(engine_cls, grid_config1):
    """The matrix sweep restricted to the targets' ancestors: targets that are
    ancestors of other targets, duplicates, top- and bottom-region targets and
    a whole root path, against the oracle's full sweep."""
    g = grid_config1
    rng = np.random.default_rng(99)
    deep = [int(x) for x in rng.integers(0, g.n // 4, 3)]
    t = []
    for r in deep:
        t += root_path(g, r)
    t += [int(x) for x in rng.integers(0, g.n, 40)] + t[:5]
    t = np.array(t, dtype=np.uint32)
    s = rng.integers(0, g.n, 70).astype(np.uint32)
    say("target paths T", len(t))
    o = Oracle(g)
    with engine_cls(g) as e:
        got, st = e.compute_matrix_flat(s, t)
        assert st.nodes_swept < g.n  # the restricted sweep ran
    assert np.array_equal(got, o.compute_matrix_flat_mt(s, t))


@pytest.mark.parametrize("lanes", [64, 32])
def test_matrix_long_paths(engine_cls, grid_config1, monkeypatch, lanes):
    """A bottom region that is the whole elimination tree: bottom paths longer
    than 64 elements and elements with more than 32 in-arcs."""
    g = grid_config1
    monkeypatch.setenv("BFG_BOTTOM_LIMIT", str(1 << 29))
    rng = np.random.default_rng(7)
    t = np.array(root_path(g, int(rng.integers(0, g.n // 8))) + [int(x) for x in rng.integers(0, g.n, 60)],
                 dtype=np.uint32)
    s = rng.integers(0, g.n, 80).astype(np.uint32)
    say("long paths lanes", lanes, "T", len(t))
    with engine_cls(g) as e:
        e.configure(lanes=lanes)
        info = e.info()
        assert info["n_top"] == 0
        got, st = e.compute_matrix_flat(s, t)
    assert np.array_equal(got, Oracle(g).compute_matrix_flat_mt(s, t))


@pytest.mark.parametrize("lanes", [64, 16])
def test_matrix_scalar_intra_blocks(engine_cls, grid_config1, monkeypatch, lanes):
    """Every DOWN level relaxes its intra-segment blocks with scalar-loaded
    weights (the form wide levels use), against the oracle."""
    g = grid_config1
    monkeypatch.setenv("BFG_TRI_SCALAR_SEGS", "0")
    rng = np.random.default_rng(11)
    s = rng.integers(0, g.n, 90).astype(np.uint32)
    t = rng.integers(0, g.n, 70).astype(np.uint32)
    say("scalar intra blocks lanes", lanes)
    with engine_cls(g) as e:
        e.configure(lanes=lanes)
        got, _ = e.compute_matrix_flat(s, t)
        rows = e.query_batch(s[:20]).dist
    o = Oracle(g)
    assert np.array_equal(got, o.compute_matrix_flat_mt(s, t))
    all_nodes = np.arange(g.n, dtype=np.uint32)
    assert np.array_equal(rows.reshape(-1), o.compute_matrix_flat_mt(s[:20], all_nodes))


@pytest.mark.parametrize("lanes,cap", [(32, 0), (16, 48), (64, 64)])
def test_matrix_lanes_and_batches(engine_cls, grid_config1, lanes, cap):
    """K in {16, 32, 64}, and a lane capacity that forces several device batches."""
    g = grid_config1
    say("lanes", lanes, "cap", cap)
    rng = np.random.default_rng(lanes)
    s = rng.integers(0, g.n, 200).astype(np.uint32)
    t = rng.integers(0, g.n, 333).astype(np.uint32)
    o = Oracle(g)
    exp = o.compute_matrix_flat_mt(s, t)
    with engine_cls(g) as e:
        e.configure(lanes=lanes, lane_capacity=cap)
        got, st = e.compute_matrix_flat(s, t)
        assert np.array_equal(got, exp)
        if cap:
            eff = -(-cap // 64) * 64  # capacities round up to whole wavefronts of lanes
            assert e.info()["lane_capacity"] == eff
            assert st.batches == (200 + eff - 1) // eff


def test_query_batch_full_rows_and_edges(engine_cls, grid_small):
    g = grid_small
    o = Oracle(g)
    say("full rows n", g.n)
    src = np.array([0, g.n - 1, 5, 5, 17], dtype=np.uint32)  # extremes and a duplicate
    with engine_cls(g) as e:
        r = e.query_batch(src)
        for i, s in enumerate(src):
            assert np.array_equal(r.dist[i], o.phast_query(int(s)))
        # empty inputs
        m, _ = e.compute_matrix_flat([], [1, 2])
        assert m.size == 0
        m, _ = e.compute_matrix_flat([1, 2], [])
        assert m.size == 0
        # out-of-range rank is an error, not a crash
        from bfgpu import BfgError
        with pytest.raises(BfgError):
            e.compute_matrix_flat([g.n], [0])


def test_bounded_and_isochrones_uneven_limits(engine_cls, grid_config1):
    g = grid_config1
    o = Oracle(g)
    rng = np.random.default_rng(99)
    L = 40
    org = rng.integers(0, g.n, L).astype(np.uint32)
    lim = rng.choice([300, 600, 900, 1800, 2700, 3600], L).astype(np.uint32)
    say("isochrones", L)
    with engine_cls(g) as e:
        bits, times = e.isochrones(org, lim)
        for i in range(L):
            full = o.phast_query(int(org[i]))
            exp = mask(full, lim[i])
            assert np.array_equal(times[i], exp)
            # reach bits
            reach = np.unpackbits(bits[i].view(np.uint8), bitorder="little")[:g.n].astype(bool)
            assert np.array_equal(reach, exp != INF)
            # same lane, reference bounded engines (values > T masked)
        blk = e.query_batch_block_gated(org[:8], 900)
        ref = o.query_batch_block_gated(org[:8], 900)
        assert np.array_equal(blk.dist, mask(ref, 900))


def test_seeded_multi_seed(engine_cls, grid_small):
    g = grid_small
    o = Oracle(g)
    rng = np.random.default_rng(4)
    lists, thr = [], []
    for i in range(20):
        k = int(rng.integers(1, 6))
        lists.append([(int(r), int(c)) for r, c in zip(rng.integers(0, g.n, k), rng.integers(0, 200, k))])
        thr.append(int(rng.integers(50, 3000)))
    lists.append([(3, 10), (3, 4)])  # duplicate seed rank keeps the smaller cost
    thr.append(5000)
    say("seeded", len(lists))
    with engine_cls(g) as e:
        rows = e.phast_seeded(lists, thr)
        for i, seeds in enumerate(lists):
            exp = np.full(g.n, INF, dtype=np.uint32)
            for r, d in o.seeded(seeds, thr[i]):
                exp[r] = d
            assert np.array_equal(rows[i], exp), i
        got = e.run_phast_bounded_fast_seeded(lists[0], thr[0])
        assert got == o.seeded(lists[0], thr[0])


def test_chords_graph(engine_cls):
    """A grid with long motorway chords (the country-graph generator at small size)."""
    g = grid_graph(64, chord_frac=0.02, seed=7)
    o = Oracle(g)
    rng = np.random.default_rng(8)
    s = rng.integers(0, g.n, 64).astype(np.uint32)
    t = rng.integers(0, g.n, 64).astype(np.uint32)
    say("chords n", g.n)
    with engine_cls(g) as e:
        got, _ = e.compute_matrix_flat(s, t)
    assert np.array_equal(got, o.compute_matrix_flat(s, t))


@pytest.mark.parametrize("lanes,thr", [(64, INF), (32, 2500)])
def test_table_seeded(engine_cls, grid_config1, lanes, thr):
    """Seeded tables (table_phast_lopsided): multi-seed sources and targets with
    parts, empty seed sets, targets sharing a seed rank with a source
    (conflicts), unbounded and bounded, against the oracle."""
    g = grid_config1
    rng = np.random.default_rng(lanes + 1)
    ranks = rng.choice(g.n, 400, replace=False)

    def sets(count, pool):
        out = []
        for _ in range(count):
            k = int(rng.integers(0, 4))
            out.append([(int(r), int(rng.integers(0, 400)), 0, True) for r in rng.choice(pool, k, replace=False)])
        return out

    src = sets(70, ranks[:200])
    tgt = sets(90, ranks[150:])  # ranks 150..199 may be shared: conflicts
    say("seeded table lanes", lanes, "thr", thr)
    with engine_cls(g) as e:
        e.configure(lanes=lanes)
        cells, conf = e.table_seeded_bounded_routed(src, tgt, thr)
    exp, econf = Oracle(g).table_phast_lopsided(src, tgt, thr)
    assert conf.any() and (exp != INF).any()
    assert np.array_equal(conf, econf)
    assert np.array_equal(cells, exp)


def test_reverse_phast_and_arrive_isochrones(engine_cls, grid_config1):
    """Reverse PHAST (range/phast.rs:879-1108) through the engine over the
    transposed CCH: all-to-one rows, bounded, and the reverse matrix equals the
    forward one transposed."""
    g = grid_config1
    rng = np.random.default_rng(3)
    tg = rng.integers(0, g.n, 12).astype(np.uint32)
    o = Oracle(g)
    say("reverse PHAST targets", len(tg))
    with engine_cls(g) as e:
        r = e.query_block_gated_reverse(tg).dist
        rb = e.query_block_gated_reverse(tg, 1500).dist
        s = rng.integers(0, g.n, 30).astype(np.uint32)
        fwd, _ = e.compute_matrix_flat(s, tg)
        rev, _ = e.reverse_engine().compute_matrix_flat(tg, s)
        lim = np.array([600, 1200, 2400, 4800] * 3, dtype=np.uint32)
        bits, times = e.reverse_engine().isochrones(tg, lim)
    for i, t in enumerate(tg):
        exp = o.phast_query_reverse(int(t))
        assert np.array_equal(r[i], exp)
        assert np.array_equal(rb[i], mask(exp, 1500))
        assert np.array_equal(times[i], mask(exp, lim[i]))  # arrive isochrone: who reaches t within the limit
        inside = np.nonzero(exp <= lim[i])[0]
        assert np.array_equal(np.nonzero(np.unpackbits(bits[i].view(np.uint8), bitorder="little")[:g.n])[0], inside)
    assert np.array_equal(rev.reshape(tg.size, s.size).T, fwd.reshape(s.size, tg.size))


def test_autotune_and_tuning_keep_results(engine_cls, grid_config1):
    """The autotuner's candidates (label width, K, DOWN unit size, source
    tile) and the scalar-block threshold change timings only: the matrix
    stays bit-exact."""
    from bfgpu.autotune import autotune_matrix
    from bfgpu.hip import DeviceArray
    g = grid_config1
    rng = np.random.default_rng(12)
    s = rng.integers(0, g.n, 256).astype(np.uint32)
    t = rng.integers(0, g.n, 120).astype(np.uint32)
    exp = Oracle(g).compute_matrix_flat_mt(s, t)
    say("autotune 256x120")
    with engine_cls(g) as e:
        ds, dt = DeviceArray.from_host(s), DeviceArray.from_host(t)
        dout = DeviceArray(s.size * t.size, np.uint32)
        knobs = (("label_bits", (16, 32)), ("lanes", (64, 16)), ("plan_rows", (64, 512)), ("source_tile", (1, 2)))
        best, timings = autotune_matrix(e, ds.ptr, s.size, dt.ptr, t.size, dout.ptr, reps=1, knobs=knobs)
        assert len(timings) == 5 and e.info()["lanes"] == best["lanes"]
        assert np.array_equal(dout.to_host(), exp)
        for src_tile in (2, 4):  # several sweeps per matrix, each a tile of the sources
            cap = (s.size + src_tile - 1) // src_tile
            e.configure(lane_capacity=cap)
            got, st = e.compute_matrix_flat(s, t)
            per_batch = (cap + 63) // 64 * 64  # capacities round up to whole wavefronts of lanes
            assert np.array_equal(got, exp) and st.batches == (s.size + per_batch - 1) // per_batch
        e.configure(lane_capacity=0)
        e.tune(plan_rows=64, tri_scalar_segs=0)
        got, _ = e.compute_matrix_flat(s, t)
        assert np.array_equal(got, exp)


def test_isochrone_lists_against_seeded_oracle(engine_cls, grid_config1):
    """bfg_isochrone_lists: per lane the settled (rank, dist <= limit) pairs in
    ascending rank, exactly run_phast_bounded_fast_seeded's output, for
    single- and multi-seed lanes and limits from 0 to INF; the reach bits
    mark exactly the listed ranks."""
    g = grid_config1
    o = Oracle(g)
    rng = np.random.default_rng(31)
    L = 150
    lanes = [[(int(r), int(c))] for r, c in zip(rng.integers(0, g.n, L), rng.integers(0, 50, L))]
    for i in range(0, L, 7):
        lanes[i] = lanes[i] + [(int(rng.integers(0, g.n)), int(rng.integers(0, 200)))]
    thr = rng.choice(np.array([0, 60, 300, 900, 1800, 3600, INF], dtype=np.uint32), L)
    say("isochrone lists", L)
    with engine_cls(g) as e:
        lists, bits = e.isochrone_lists(lanes, thr, want_bits=True)
        small = e.isochrone_lists(lanes[:10], thr[:10], capacity=5)  # too small: counted, then rerun
    words = (g.n + 63) // 64
    for i in range(L):
        exp = o.seeded(lanes[i], int(thr[i]))
        assert lists[i].tolist() == [list(x) for x in exp], i
        inside = np.zeros(words * 64, dtype=bool)
        inside[lists[i][:, 0]] = True
        assert np.array_equal(bits[i], np.packbits(inside, bitorder="little").view(np.uint64))
    for i in range(10):
        assert np.array_equal(small[i], lists[i])


def test_isochrone_lists_buffer_too_small_reports_sizes(engine_cls, grid_small):
    """A pair buffer shorter than the result: BFG_BUFFER_TOO_SMALL with the
    sizes filled and nothing written past the capacity."""
    import ctypes as C
    from bfgpu import _native
    from bfgpu.engine import BFG_BUFFER_TOO_SMALL
    g = grid_small
    rng = np.random.default_rng(32)
    L = 12
    org = rng.integers(0, g.n, L).astype(np.uint32)
    off = np.arange(L + 1, dtype=np.uint32)
    cost = np.zeros(L, dtype=np.uint32)
    thr = np.full(L, 600, dtype=np.uint32)
    start = np.zeros(L, dtype=np.uint64)
    length = np.zeros(L, dtype=np.uint32)
    total = np.zeros(1, dtype=np.uint64)
    pairs = np.full(2 * 8 + 2, 0xABCDEF, dtype=np.uint32)
    P = _native.ptr
    with engine_cls(g) as e:
        rc = e._lib.bfg_isochrone_lists(e._h, P(off), P(org), P(cost), L, P(thr), None, P(start, _native.u64p),
                                        P(length), P(pairs), 8, P(total, _native.u64p))
        want = e.isochrone_lists([[(int(r), 0)] for r in org], thr)
    assert rc == BFG_BUFFER_TOO_SMALL
    assert int(total[0]) == sum(len(x) for x in want) > 8
    assert length.tolist() == [len(x) for x in want]
    assert (pairs[16:] == 0xABCDEF).all()


def test_concurrent_calls_equal_sequential(engine_cls, grid_config1):
    """Calls from several threads on one context run at once (each on its own
    stream and arena) and give the same results as one after another."""
    import threading
    g = grid_config1
    rng = np.random.default_rng(33)
    jobs = []
    for k in range(12):
        if k % 3 == 0:
            jobs.append(("iso", rng.integers(0, g.n, 40).astype(np.uint32),
                         rng.choice(np.array([300, 900, 1800], dtype=np.uint32), 40)))
        else:
            jobs.append(("mat", rng.integers(0, g.n, int(rng.integers(5, 200))).astype(np.uint32),
                         rng.integers(0, g.n, int(rng.integers(5, 300))).astype(np.uint32)))
    say("concurrent calls", len(jobs))
    with engine_cls(g) as e:
        e.set_concurrency(4)

        def run(j):
            kind, a, b = j
            if kind == "iso":
                return e.isochrones(a, b, want_times=False)[0]
            return e.compute_matrix_flat(a, b)[0]
        want = [run(j) for j in jobs]
        got = [None] * len(jobs)
        errs = []

        def worker(idx):
            try:
                for i in idx:
                    got[i] = run(jobs[i])
            except Exception as ex:  # surfaced below
                errs.append(ex)
        ts = [threading.Thread(target=worker, args=(list(range(t, len(jobs), 4)),)) for t in range(4)]
        for t in ts:
            t.start()
        for t in ts:
            t.join()
    assert not errs, errs
    for a, b in zip(got, want):
        assert np.array_equal(a, b)


def test_gate_skips_and_stays_exact(engine_cls, grid_config1):
    """Short limits: the block gate leaves most (segment, wave group) units and
    bottom trees closed, and every lane still equals the oracle's seeded
    PHAST; a wave group holding one unbounded lane keeps everything open."""
    g = grid_config1
    o = Oracle(g)
    rng = np.random.default_rng(34)
    L = 192
    org = rng.integers(0, g.n, L)
    lanes = [[(int(r), 0)] for r in org]
    thr = np.full(L, 120, dtype=np.uint32)
    thr[150] = INF  # the third wave group holds an unbounded lane
    say("gate", L)
    with engine_cls(g) as e:
        lists = e.isochrone_lists(lanes, thr)
        st = e.last_stats(L)
    say("gate open", st.gate_seg_open, "of", st.gate_seg_units, "segments;", st.gate_bt_open, "of",
        st.gate_bt_units, "trees")
    nseg = st.gate_seg_units // 3
    assert st.gate_seg_open < 0.5 * 2 * nseg + nseg  # two wave groups mostly closed, one open
    assert st.gate_bt_open < st.gate_bt_units
    for i in range(L):
        assert lists[i].tolist() == [list(x) for x in o.seeded(lanes[i], int(thr[i]))], i


def test_matrix_multi_equals_requests_alone(engine_cls, grid_config1):
    """bfg_matrix_multi: requests of uneven shapes and thresholds share one
    sweep; each block equals the request issued alone and the oracle, also
    when the lanes span several batches."""
    g = grid_config1
    o = Oracle(g)
    rng = np.random.default_rng(35)
    reqs = []
    for k in range(9):
        S, T = int(rng.integers(1, 90)), int(rng.integers(1, 120))
        thr = [INF, 900, 2400][k % 3]
        reqs.append((rng.integers(0, g.n, S).astype(np.uint32), rng.integers(0, g.n, T).astype(np.uint32), thr))
    reqs.append((reqs[0][0][:3], reqs[0][1], INF))  # a request that repeats another's targets
    say("matrix multi", len(reqs))
    with engine_cls(g) as e:
        got = e.compute_matrix_multi(reqs)
        alone = [e.compute_matrix_block_gated(s, t, thr)[0] for s, t, thr in reqs]
        e.configure(lane_capacity=128)  # several batches; requests cross batch boundaries
        got2 = e.compute_matrix_multi(reqs)
        e.configure(lane_capacity=0)
    for (s, t, thr), a, b, c in zip(reqs, got, alone, got2):
        assert np.array_equal(a, b)
        assert np.array_equal(c, b)
        assert np.array_equal(a, mask(o.compute_matrix_flat_mt(s, t), thr))


@pytest.mark.parametrize("scale", [1, 40])
def test_label_width_16_equals_32(engine_cls, grid_config1, scale):
    """The packed 16-bit DOWN rows give the u32 rows' results bit for bit: on
    the grid as it is (labels fit 16 bits), and with every weight times 40 so
    that labels pass 0xFFFF and units fall back to 32 bits; both against the
    oracle, unbounded and bounded, K = 64 and 16."""
    from bfgpu import CchGraph
    g0 = grid_config1
    g = g0
    if scale != 1:
        def sc(w):
            w = w.astype(np.uint64) * scale
            return np.where(w >= INF, INF, w).astype(np.uint32)
        g = CchGraph(n=g0.n, up_off=g0.up_off, up_tgt=g0.up_tgt, up_w=sc(g0.up_w),
                     dn_off=g0.dn_off, dn_tgt=g0.dn_tgt, dn_w=sc(g0.dn_w))
    rng = np.random.default_rng(36 + scale)
    s = rng.integers(0, g.n, 200).astype(np.uint32)
    t = rng.integers(0, g.n, 150).astype(np.uint32)
    o = Oracle(g)
    exp = o.compute_matrix_flat_mt(s, t)
    if scale != 1:
        assert exp[exp != INF].max() > 0xFFFF  # the fallback is exercised
    say("label width", scale)
    with engine_cls(g) as e:
        for lanes in (64, 16):
            e.configure(lanes=lanes)
            res = {}
            for bits in (16, 32):
                e.set_label_width(bits)
                res[bits] = e.compute_matrix_flat(s, t)[0]
                assert np.array_equal(res[bits], exp), (lanes, bits)
                mb = e.compute_matrix_block_gated(s, t, 20000 * scale)[0]
                assert np.array_equal(mb, mask(exp, 20000 * scale)), (lanes, bits)
        e.configure(lanes=64)
        e.set_label_width(16)
