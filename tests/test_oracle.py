"""Pins the CPU oracle before it is trusted as the GPU checker.

Each check is chosen so that a plausible slip in the restatement (a dropped
term, a wrong index, a transposed operand, a missing INF filter) fails:
  * the reference's own hand vectors (range/phast.rs:1198-1554,
    bucket_ch.rs:4860-4908);
  * the broom closed form d(s->t) = wu[s] + wd[t] (bucket_ch.rs:5638-5670);
  * brute-force Dijkstra on the CCH union graph and on the edge-based graph;
  * the invariants the reference asserts in its bench and validators:
    K-lane == single source (bench/main.rs:1392-1434), SoA == AoS,
    bounded == unbounded within T (bucket_ch.rs:5672-5736), gated ==
    ungated within T (range/mod.rs:818-951), seeds == super-source.
"""
import numpy as np
import pytest

from conftest import broom_expect, broom_graph, hand_graph
from oracle.oracle import INF, Oracle, dijkstra, union_graph


def mask(d, thr):
    d = np.array(d, dtype=np.uint32, copy=True)
    d[d > thr] = INF
    return d


@pytest.mark.parametrize("name", ["five_node", "asymmetric_three_node"])
def test_hand_vectors_forward(name):
    g, hv = hand_graph(name)
    o = Oracle(g)
    for s, exp in hv["forward"].items():
        assert o.phast_query(int(s)).tolist() == exp
        assert o.query_batch([int(s)])[0].tolist() == exp
        assert o.query_batch_soa([int(s)])[0].tolist() == exp


@pytest.mark.parametrize("name", ["five_node", "asymmetric_three_node"])
def test_hand_vectors_reverse_and_transpose(name):
    g, hv = hand_graph(name)
    o = Oracle(g)
    for t, exp in hv["reverse"].items():
        assert o.phast_query_reverse(int(t)).tolist() == exp
    n = g.n
    fwd = np.stack([o.phast_query(s) for s in range(n)])
    rev = np.stack([o.phast_query_reverse(t) for t in range(n)])
    assert np.array_equal(fwd, rev.T)  # d_fwd(s)[t] == d_rev(t)[s], phast.rs:1371-1398


def test_hand_vector_bounded_reverse():
    g, hv = hand_graph("five_node")
    b = hv["bounded_reverse"]
    d = Oracle(g).phast_query_reverse(b["target"])
    within = {i: int(x) for i, x in enumerate(d) if x <= b["threshold"]}
    assert len(within) == b["n_reachable"]
    assert within == {int(k): v for k, v in b["within"].items()}


def test_asymmetric_matrix():
    g, hv = hand_graph("asymmetric_three_node")
    o = Oracle(g)
    m = o.compute_matrix_flat([0, 1, 2], [0, 1, 2]).reshape(3, 3)
    exp = np.array([hv["forward"][str(s)] for s in range(3)], dtype=np.uint32)
    assert np.array_equal(m, exp)
    assert m[0, 1] != m[1, 0]


def test_inf_arc_flats():
    g, hv = hand_graph("inf_arc_four_node")
    o = Oracle(g)
    uo, ut, uw = o.up_flat()
    assert uo.tolist() == hv["up_flat"]["offsets"]
    assert ut.tolist() == hv["up_flat"]["targets"] and uw.tolist() == hv["up_flat"]["weights"]
    do, dt, dw = o.down_flat()
    assert do.tolist() == hv["down_flat"]["offsets"]
    assert dt.tolist() == hv["down_flat"]["targets"] and dw.tolist() == hv["down_flat"]["weights"]
    ro, rs, rw = o.down_reverse_flat()
    assert ro.tolist() == hv["down_reverse_flat"]["offsets"]
    assert rs.tolist() == hv["down_reverse_flat"]["sources"] and rw.tolist() == hv["down_reverse_flat"]["weights"]
    # the INF parallel arcs never win: PHAST equals Dijkstra on the union graph
    off, head, w = union_graph(g)
    for s in range(g.n):
        assert np.array_equal(o.phast_query(s), dijkstra(g.n, off, head, w, s))


@pytest.mark.parametrize("n_leaf", [5, 12, 300])
def test_broom_closed_form(n_leaf):
    g, wu, wd = broom_graph(n_leaf)
    o = Oracle(g)
    s = np.arange(n_leaf)
    m = o.compute_matrix_flat(s, s).reshape(n_leaf, n_leaf)
    assert np.array_equal(m, broom_expect(wu, wd, s, s))
    m2 = o.compute_matrix_flat_soa(s, s).reshape(n_leaf, n_leaf)
    assert np.array_equal(m2, m)
    for thr in (0, 15, 25, 55, 1000):
        mb = o.compute_matrix_block_gated(s, s, thr).reshape(n_leaf, n_leaf)
        assert np.array_equal(mask(mb, thr), mask(m, thr))


def test_grid_matches_bruteforce(grid_small):
    """PHAST == Dijkstra on the CCH union graph == Dijkstra on the EBG."""
    g = grid_small
    o = Oracle(g)
    off, head, w = union_graph(g)
    rng = np.random.default_rng(7)
    for s in rng.integers(0, g.n, 12):
        p = o.phast_query(int(s))
        assert np.array_equal(p, dijkstra(g.n, off, head, w, int(s)))
        assert np.array_equal(p, dijkstra(g.n, g.ebg_off, g.ebg_head, g.ebg_w, int(s)))


def test_klane_equals_single_source(grid_small):
    g = grid_small
    o = Oracle(g)
    rng = np.random.default_rng(11)
    src = rng.integers(0, g.n, 8)
    kb = o.query_batch(src)
    soa = o.query_batch_soa(src)
    assert np.array_equal(kb, soa)
    for i, s in enumerate(src):
        assert np.array_equal(kb[i], o.phast_query(int(s)))
    tgt = rng.integers(0, g.n, 50)
    m = o.compute_matrix_flat(src, tgt).reshape(8, 50)
    assert np.array_equal(m, kb[:, tgt])
    assert np.array_equal(o.compute_matrix_flat_mt(src, tgt, 4), m.reshape(-1))


def test_bounded_and_gated_equal_unbounded_within_threshold(grid_config1):
    g = grid_config1
    assert g.n >= 4 * 4096  # exercises the real block-gated path, not its small-graph fallback
    o = Oracle(g)
    rng = np.random.default_rng(3)
    src = rng.integers(0, g.n, 8)
    full = o.query_batch(src)
    for thr in (60, 300, 900, 1800):
        assert np.array_equal(mask(o.query_batch_bounded(src, thr), thr), mask(full, thr))
        assert np.array_equal(mask(o.query_batch_block_gated(src, thr), thr), mask(full, thr))


def test_seeded_is_a_super_source(grid_small):
    g = grid_small
    o = Oracle(g)
    rng = np.random.default_rng(5)
    for _ in range(5):
        k = int(rng.integers(1, 5))
        seeds = [(int(r), int(c)) for r, c in zip(rng.integers(0, g.n, k), rng.integers(0, 120, k))]
        thr = int(rng.integers(100, 2000))
        got = dict(o.seeded(seeds, thr))
        best = np.full(g.n, INF, dtype=np.uint64)
        for r, c in seeds:
            d = o.phast_query(r).astype(np.uint64)
            d = np.where(d == INF, INF, d + c)
            best = np.minimum(best, d)
        exp = {int(v): int(best[v]) for v in range(g.n) if best[v] <= thr}
        assert got == exp
        assert list(got) == sorted(got)  # ascending rank (isochrone_handler.rs:456-471)


def test_saturating_add_at_inf_boundary():
    """Weights near u32::MAX saturate instead of wrapping (Rust saturating_add)."""
    from bfgpu import CchGraph
    big = 0xFFFFFFF0
    g = CchGraph.from_arcs(3, [(0, 2, big), (1, 2, 5)], [(2, 0, 100), (2, 1, 0xFFFFFFFE)])
    o = Oracle(g)
    d = o.phast_query(0)
    assert d.tolist() == [0, INF, big]  # big + 0xFFFFFFFE saturates to INF
    d1 = o.phast_query(1)
    assert d1.tolist() == [105, 0, 5]


def _seedsets(rng, ranks, count, max_seeds, max_part, allow_empty=True):
    out = []
    for _ in range(count):
        k = int(rng.integers(0 if allow_empty else 1, max_seeds + 1))
        out.append([(int(r), int(rng.integers(0, max_part + 1)), 0, True)
                    for r in rng.choice(ranks, k, replace=False)])
    return out


@pytest.mark.parametrize("thr", [INF, 0, 60, 150])
def test_table_phast_lopsided_broom_closed_form(thr):
    """The seeded table on the broom (bucket_ch.rs:5638-5670): a source's field
    at leaf t is min over its seeds (s, a) of a + wu[s] + wd[t] (leaves differ),
    present up to threshold + its largest part + the largest target part; a
    cell is min over the target's seeds (t, b) of field - b (saturating), INF
    above threshold; a shared seed rank is a conflict."""
    g, wu, wd = broom_graph(12)
    rng = np.random.default_rng(thr & 0xFFFF)
    leaves = np.arange(12)
    src = _seedsets(rng, leaves, 9, 3, 40)
    tgt = _seedsets(rng, leaves, 11, 3, 40)
    cells, conf = Oracle(g).table_phast_lopsided(src, tgt, thr)
    max_b = max((s[1] for v in tgt for s in v), default=0)
    for i, ss in enumerate(src):
        bound = INF if thr == INF else min(INF, thr + max((s[1] for s in ss), default=0) + max_b)
        for j, ts in enumerate(tgt):
            k = i * len(tgt) + j
            shared = {s[0] for s in ss} & {t[0] for t in ts}
            assert conf[k] == (1 if ss and ts and shared else 0)
            if not ss or not ts or shared:
                assert cells[k] == INF
                continue
            best = INF
            for t, b, _, _ in ts:
                f = min(int(a + wu[s] + wd[t]) for s, a, _, _ in ss)
                if f <= bound:
                    best = min(best, max(0, f - b))
            if thr != INF and best > thr:
                best = INF
            assert cells[k] == best, (i, j)


def test_table_phast_lopsided_single_seeds_equal_the_matrix(grid_small):
    """One seed of part 0 per source and per target, no shared rank: the table
    is compute_matrix_flat."""
    g = grid_small
    rng = np.random.default_rng(5)
    ranks = rng.choice(g.n, 60, replace=False)
    s, t = ranks[:25], ranks[25:]
    cells, conf = Oracle(g).table_phast_lopsided([[(int(r), 0)] for r in s], [[(int(r), 0)] for r in t], INF)
    assert not conf.any()
    assert np.array_equal(cells, Oracle(g).compute_matrix_flat(s, t))


def test_transposed_phast_is_reverse_phast(grid_small):
    """Forward PHAST on the transposed CCH (CchGraph.transpose) equals the
    reverse PHAST restatement (range/phast.rs:879-1108) and Dijkstra towards
    the target on the transposed union graph."""
    g = grid_small
    o, ro = Oracle(g), Oracle(g.transpose())
    off, head, w = union_graph(g.transpose())
    for t in (0, 7, g.n // 2, g.n - 1):
        rev = o.phast_query_reverse(t)
        assert np.array_equal(ro.phast_query(t), rev)
        assert np.array_equal(rev, dijkstra(g.n, off, head, w, t))


@pytest.mark.parametrize("n_leaf,thresholds", [(5, (0, 15, 25, 55, 1000)), (12, (0, 30, 75, 130, 100_000))])
def test_table_bucket_broom_bounded(n_leaf, thresholds):
    """bucket_ch.rs:5617-5736 (max_minutes_bound_tests): the bucket matrix
    on the broom is the closed form; bounded, every cell within the
    threshold is exact and every other cell is above it and never below the
    unbounded one."""
    g, wu, wd = broom_graph(n_leaf)
    o = Oracle(g)
    s = np.arange(n_leaf)
    unb = o.table_bucket(s, s).reshape(n_leaf, n_leaf)
    assert np.array_equal(unb, broom_expect(wu, wd, s, s))
    for thr in thresholds:
        b = o.table_bucket(s, s, thr).reshape(n_leaf, n_leaf)
        inb = unb <= thr
        assert np.array_equal(b[inb], unb[inb])
        assert np.all(b[~inb] > thr) and np.all(b[~inb] >= unb[~inb])


@pytest.mark.parametrize("name", ["five_node", "asymmetric_three_node", "inf_arc_four_node"])
def test_table_bucket_equals_phast_on_hand_graphs(name):
    """The bucket join (forward UP sweeps, backward DOWN-reverse sweeps) and
    the PHAST matrix give the same cells: both are the least up-down path."""
    g, _ = hand_graph(name)
    o = Oracle(g)
    a = np.arange(g.n)
    assert np.array_equal(o.table_bucket(a, a), o.compute_matrix_flat(a, a))


def test_table_bucket_equals_phast_and_bruteforce_on_a_grid(grid_small):
    g = grid_small
    o = Oracle(g)
    rng = np.random.default_rng(21)
    s = rng.integers(0, g.n, 40)
    t = rng.integers(0, g.n, 30)
    t[:5] = s[:5]  # diagonal cells
    b = o.table_bucket(s, t)
    assert np.array_equal(b, o.compute_matrix_flat(s, t))
    for i in range(3):  # and Dijkstra on the edge-based graph
        d = dijkstra(g.n, g.ebg_off, g.ebg_head, g.ebg_w, int(s[i]))
        assert np.array_equal(b.reshape(s.size, t.size)[i], d[t])
    for thr in (0, 500, 2000):
        bb = o.table_bucket(s, t, thr)
        assert np.array_equal(mask(bb, thr), mask(b, thr))
