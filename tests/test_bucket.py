"""GPU parity of the bucket many-to-many (bfg_table_bucket): the forward
upward phase of the sources on the CCH and the backward upward phase of the
targets on the transposed CCH, joined at their common elimination-tree
ancestors (table_bucket_parallel_bounded, route/src/matrix/bucket_ch.rs:
3436-3520).

Bit-exact against the oracle's bucket restatement and its PHAST matrix
(u32 integer work); bounded calls against the masked unbounded cells.  Cases:
the hand graphs, the broom closed form, the 224 x 224 grid with diagonal and
same-bottom-tree pairs (sources and targets that meet below the top region),
thresholds, lane widths 16/32/64, several source and target batches, the
device entry point, the chord grid, label widths 16 and 32 (the packed
16-bit upward pushes and their u32 fallback at x40 weights), and the 16-bit
join of matrices of more than 2048 tiles against the u32 join, with its
tiles past 16 bits recomputed at 32 bits."""
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


def near_pairs(g, rng, count):
    """Sources and targets that share a bottom tree: each source with one
    of its UP neighbours (an ancestor) and with itself."""
    s = rng.integers(0, g.n, count).astype(np.uint32)
    t = []
    for v in s:
        a, b = int(g.up_off[v]), int(g.up_off[v + 1])
        t.append(int(g.up_tgt[a + rng.integers(0, b - a)]) if b > a else int(v))
    return s, np.concatenate([np.asarray(t, dtype=np.uint32), s])


@pytest.mark.parametrize("name", ["five_node", "asymmetric_three_node", "inf_arc_four_node"])
def test_bucket_hand_vectors(engine_cls, name):
    say("bucket hand", name)
    g, _ = hand_graph(name)
    o = Oracle(g)
    a = np.arange(g.n, dtype=np.uint32)
    with engine_cls(g) as e:
        m, _ = e.table_bucket_parallel(a, a)
    assert np.array_equal(m, o.table_bucket(a, a))
    assert np.array_equal(m, o.compute_matrix_flat(a, a))


@pytest.mark.parametrize("n_leaf", [12, 200_000])
def test_bucket_broom_closed_form(engine_cls, n_leaf):
    say("bucket broom", n_leaf)
    g, wu, wd = broom_graph(n_leaf)
    rng = np.random.default_rng(n_leaf + 1)
    s = rng.choice(n_leaf, min(n_leaf, 300), replace=False)
    t = rng.choice(n_leaf, min(n_leaf, 500), replace=False)
    t[0] = s[0]
    exp = broom_expect(wu, wd, s, t)
    with engine_cls(g) as e:
        m, _ = e.table_bucket_parallel(s, t)
        assert np.array_equal(m.reshape(s.size, t.size), exp)
        for thr in (0, 30, 75, 130):
            mb, _ = e.table_bucket_parallel_bounded(s, t, thr)
            assert np.array_equal(mb.reshape(s.size, t.size), mask(exp, thr))


def test_bucket_config1_grid(engine_cls, grid_config1):
    """224 x 224 grid: random pairs, diagonal cells and pairs that meet inside
    a bottom tree; unbounded and bounded; equal to the PHAST matrix."""
    g = grid_config1
    o = Oracle(g)
    rng = np.random.default_rng(31)
    s = rng.integers(0, g.n, 100).astype(np.uint32)
    t = rng.integers(0, g.n, 100).astype(np.uint32)
    ns, nt = near_pairs(g, rng, 60)
    s, t = np.concatenate([s, ns]), np.concatenate([t, nt])
    say("bucket 224 grid", s.size, "x", t.size)
    exp = o.compute_matrix_flat_mt(s, t)
    assert np.array_equal(o.table_bucket(s[:20], t), exp.reshape(s.size, t.size)[:20].reshape(-1))
    with engine_cls(g) as e:
        m, st = e.table_bucket_parallel(s, t)
        assert np.array_equal(m, exp)
        assert st.nodes_swept > 0  # some pairs met inside a bottom tree
        ph, _ = e.compute_matrix_flat(s, t)
        assert np.array_equal(m, ph)
        for thr in (0, 300, 1800):
            mb, _ = e.table_bucket_parallel_bounded(s, t, thr)
            assert np.array_equal(mb, mask(exp, thr))


@pytest.mark.parametrize("lanes,cap", [(16, 64), (32, 128), (64, 64)])
def test_bucket_lane_widths_and_batches(engine_cls, grid_config1, lanes, cap):
    """Lane groups of 16/32/64 on both sides; capacities that cut the sources
    and the targets into several batches."""
    g = grid_config1
    o = Oracle(g)
    rng = np.random.default_rng(lanes + cap)
    s = rng.integers(0, g.n, 150).astype(np.uint32)
    t = rng.integers(0, g.n, 200).astype(np.uint32)
    ns, nt = near_pairs(g, rng, 30)
    s, t = np.concatenate([s, ns]), np.concatenate([t, nt])
    say("bucket lanes", lanes, "cap", cap, s.size, "x", t.size)
    exp = o.compute_matrix_flat_mt(s, t)
    with engine_cls(g) as e:
        rev = e.pair_reverse()
        for x in (e, rev):
            x.configure(lanes=lanes, lane_capacity=cap)
        m, st = e.table_bucket_parallel(s, t)
        assert np.array_equal(m, exp)
        assert st.batches > 1


def test_bucket_device_entry_and_chords(engine_cls):
    """The device entry point on the 64 x 64 grid with long chords."""
    from bfgpu.hip import DeviceArray
    g = grid_graph(64, chord_frac=0.02)
    o = Oracle(g)
    rng = np.random.default_rng(5)
    s = rng.integers(0, g.n, 90).astype(np.uint32)
    t = rng.integers(0, g.n, 70).astype(np.uint32)
    say("bucket chords device", s.size, "x", t.size)
    exp = o.table_bucket(s, t)
    with engine_cls(g) as e:
        ds, dt = DeviceArray.from_host(s), DeviceArray.from_host(t)
        dout = DeviceArray(s.size * t.size, np.uint32)
        e.table_bucket_device(ds.ptr, s.size, dt.ptr, t.size, dout.ptr)
        assert np.array_equal(dout.to_host(), exp)
        e.table_bucket_device(ds.ptr, s.size, dt.ptr, t.size, dout.ptr, 900)
        assert np.array_equal(dout.to_host(), mask(exp, 900))


def scaled(g, k):
    """g with every finite weight times k (saturating below INF)."""
    import dataclasses
    def sc(w):
        w64 = w.astype(np.uint64) * np.uint64(k)
        return np.where(w == INF, np.uint64(INF), np.minimum(w64, np.uint64(INF - 1))).astype(np.uint32)
    return dataclasses.replace(g, up_w=sc(g.up_w), dn_w=sc(g.dn_w))


@pytest.mark.parametrize("scale", [1, 40])
@pytest.mark.parametrize("bits", [16, 32])
def test_bucket_label_widths(engine_cls, grid_config1, scale, bits):
    """The upward pushes packed 16-bit (two wave groups a wavefront) and in
    u32 give the same cells; at x40 weights labels pass 16 bits and the
    packed units fall back to the u32 rows."""
    g = grid_config1 if scale == 1 else scaled(grid_config1, scale)
    o = Oracle(g)
    rng = np.random.default_rng(7 * scale + bits)
    s = rng.integers(0, g.n, 170).astype(np.uint32)
    t = rng.integers(0, g.n, 150).astype(np.uint32)
    ns, nt = near_pairs(g, rng, 20)
    s, t = np.concatenate([s, ns]), np.concatenate([t, nt])
    say("bucket label bits", bits, "weights x", scale, s.size, "x", t.size)
    exp = o.compute_matrix_flat_mt(s, t)
    with engine_cls(g) as e:
        rev = e.pair_reverse()
        for x in (e, rev):
            x.set_label_width(bits)
        m, _ = e.table_bucket_parallel(s, t)
        assert np.array_equal(m, exp)
        mb, _ = e.table_bucket_parallel_bounded(s, t, 2000 * scale)
        assert np.array_equal(mb, mask(exp, 2000 * scale))


@pytest.mark.parametrize("scale", [1, 40])
def test_bucket_join16(engine_cls, grid_config1, scale, monkeypatch):
    """Matrices of more than 2048 tiles join one wavefront a tile on 16-bit
    label copies (packed u16 cells, labels capped at 0x7FFF): equal to the
    u32 join (BFG_JOIN_U32=1), to the PHAST matrix and, on a sample of rows,
    to the oracle.  At x40 weights most cells pass the cap and their tiles
    are recomputed at 32 bits; bounds just below, at and above the cap and
    0xFFFF."""
    g = grid_config1 if scale == 1 else scaled(grid_config1, scale)
    o = Oracle(g)
    rng = np.random.default_rng(11 + scale)
    s = rng.integers(0, g.n, 2900).astype(np.uint32)
    t = rng.integers(0, g.n, 2900).astype(np.uint32)
    ns, nt = near_pairs(g, rng, 60)
    s, t = np.concatenate([s, ns]), np.concatenate([t, nt])
    say("bucket join16 weights x", scale, s.size, "x", t.size)
    with engine_cls(g) as e:
        e.pair_reverse()
        m16, st = e.table_bucket_parallel(s, t)
        assert st.join_label_bits == 16
        wide = int(np.count_nonzero((m16 >= 0x7FFF) & (m16 != INF)))  # past the 16-bit cap
        assert (wide > 0) == (scale > 1)
        if wide:
            assert st.join_tiles_u32 > 0
        else:  # none but tiles whose join cell a pair meeting in a bottom tree lowers
            assert st.join_tiles_u32 < 50
        monkeypatch.setenv("BFG_JOIN_U32", "1")
        m32, st32 = e.table_bucket_parallel(s, t)
        assert st32.join_label_bits == 32 and st32.join_tiles_u32 == 0
        monkeypatch.delenv("BFG_JOIN_U32")
        assert np.array_equal(m16, m32)
        ph, _ = e.compute_matrix_flat(s, t)
        assert np.array_equal(m16, ph)
        exp = o.compute_matrix_flat_mt(s[:8], t)
        assert np.array_equal(m16.reshape(s.size, t.size)[:8].reshape(-1), exp)
        for thr in (1800 * scale, 0x7FFE, 0x7FFF, 0xFFFF, 100_000):
            mb, _ = e.table_bucket_parallel_bounded(s, t, thr)
            assert np.array_equal(mb, mask(ph, thr))
