"""The shape router: the restated phast_wins predicate and its cost cells
(route/src/matrix/bucket_ch.rs:6660-6711) pinned in closed form, the seeded
table's routing, and the GPU's forward/reverse lane routing of plain
matrices (equal cells either way)."""
import numpy as np
import pytest

from bfgpu import router
from bfgpu.graph import INF
from oracle.oracle import Oracle


def cells(sweep=0, scan=0):
    a, b = router.CostCell(), router.CostCell()
    a.value, b.value = sweep, scan
    return a, b


def test_ewma_update_formula():
    c = router.CostCell()
    router.update_cost_ewma(c, 1000)
    assert c.value == 1000  # the first sample as is
    router.update_cost_ewma(c, 2000)
    assert c.value == 1000 - 250 + 500
    router.update_cost_ewma(c, 0)
    assert c.value == 1250 - 312 + 0
    d = router.CostCell()
    router.update_cost_ewma(d, 0)
    assert d.value == 1  # never below 1


def test_phast_wins_structural_boundary(monkeypatch):
    monkeypatch.delenv("BUTTERFLY_MATRIX_ALGO", raising=False)
    sw, sc = cells()
    n = 400_000  # sweep cost clamp(n / 400, 500, 20000) = 1000, scan cost n
    # S * n < (S + T) * 1000  <=>  T > S * (n / 1000 - 1) = 399 S
    assert not router.phast_wins(1, 399, n, sw, sc)
    assert router.phast_wins(1, 400, n, sw, sc)
    assert not router.phast_wins(2, 798, n, sw, sc) and router.phast_wins(2, 799, n, sw, sc)
    assert not router.phast_wins(5, 4, n, sw, sc)  # tall shapes never (the reverse field is a follow-up)
    # the clamp: tiny graphs sweep at 500, huge ones at 20000
    assert router.phast_wins(1, 100_000, 1000, sw, sc) == (1000 < 100_001 * 500)
    assert router.phast_wins(1, 10, 10_000_000, sw, sc) == (10_000_000 < 11 * 20_000)


def test_phast_wins_measured_and_overrides(monkeypatch):
    monkeypatch.delenv("BUTTERFLY_MATRIX_ALGO", raising=False)
    sw, sc = cells(sweep=100, scan=1000)  # both measured: S * 1000 < (S + T) * 100  <=>  T > 9 S
    assert not router.phast_wins(10, 90, 5_000_000, sw, sc)
    assert router.phast_wins(10, 91, 5_000_000, sw, sc)
    sw, sc = cells(sweep=100, scan=0)  # one cell unmeasured: structural
    assert router.phast_wins(10, 91, 5_000_000, sw, sc) == (10 * 5_000_000 < 101 * 12_500)
    monkeypatch.setenv("BUTTERFLY_MATRIX_ALGO", "bucket")
    assert not router.phast_wins(1, 10**9, 1000)
    monkeypatch.setenv("BUTTERFLY_MATRIX_ALGO", "phast")
    assert router.phast_wins(3, 3, 10**9) and not router.phast_wins(4, 3, 10**9)


class TableStub:
    n_nodes = 400_000

    def __init__(self):
        self.calls = 0

    def table_seeded_bounded_routed(self, src, tgt, thr):
        self.calls += 1
        return np.zeros(len(src) * len(tgt), np.uint32), np.zeros(len(src) * len(tgt), np.uint8)


def test_seeded_table_routing(monkeypatch):
    monkeypatch.delenv("BUTTERFLY_MATRIX_ALGO", raising=False)
    monkeypatch.setattr(router, "SWEEP_NS", router.CostCell())
    monkeypatch.setattr(router, "SCAN_NS", router.CostCell())
    eng, used = TableStub(), []

    def bucket(s, t, thr):
        used.append((len(s), len(t)))
        return np.ones(len(s) * len(t), np.uint32), np.zeros(len(s) * len(t), np.uint8)
    src = [[(1, 0)]]
    out, _ = router.table_seeded_bounded_routed(eng, src, [[(2, 0)]] * 50, INF, bucket=bucket)
    assert used == [(1, 50)] and eng.calls == 0 and (out == 1).all()  # structural: bucket wins
    out, _ = router.table_seeded_bounded_routed(eng, src, [[(2, 0)]] * 500, INF, bucket=bucket)
    assert eng.calls == 1 and router.SCAN_NS.load() > 0  # PHAST wins; its scan time is measured
    out, _ = router.table_seeded_bounded_routed(eng, src, [[(2, 0)]] * 5, INF)  # no bucket engine: the GPU table
    assert eng.calls == 2
    out, conf = router.table_seeded_bounded_routed(eng, [], [[(2, 0)]], INF)
    assert out.size == 0 and conf.size == 0


class MatStub:
    """Forward and reverse matrices from the oracle (reverse = the transposed
    CCH's engine, whose rows are all-to-one distances)."""

    def __init__(self, g, rev=None):
        self.g, self.o, self.calls = g, Oracle(g), 0
        self._rev = rev

    def compute_matrix_block_gated(self, s, t, thr):
        self.calls += 1
        m = self.o.compute_matrix_flat(s, t)
        return (np.where(m > thr, INF, m).astype(np.uint32) if thr != INF else m), None

    def reverse_engine(self):
        if self._rev is None:
            self._rev = MatStub(self.g.transpose())
        return self._rev

    def table_bucket_parallel_bounded(self, s, t, thr):
        self.calls += 1
        m = self.o.table_bucket(s, t, thr)

        class St:  # the engine's phase times of a bucket call
            upward_time_ms, downward_time_ms = 3.0, 1.0
        return np.where(m > thr, INF, m).astype(np.uint32), St()


def test_matrix_router_sides_agree(grid_small):
    g = grid_small
    rng = np.random.default_rng(3)
    eng = MatStub(g)
    r = router.MatrixRouter(eng)
    for S, T in ((30, 5), (5, 30), (12, 12)):
        s = rng.integers(0, g.n, S).astype(np.uint32)
        t = rng.integers(0, g.n, T).astype(np.uint32)
        got = r.compute_matrix_block_gated(s, t, 900)
        exp = Oracle(g).compute_matrix_flat(s, t)
        assert np.array_equal(got, np.where(exp > 900, INF, exp))
        assert r.last_side == ("reverse" if S > T else "forward")
    assert r.fwd_ns.load() > 0 and r.rev_ns.load() > 0


def test_matrix_router_bucket_side(grid_small):
    """Square shapes go to the bucket many-to-many (tried first once both
    sides have BUCKET_MIN lanes, then by measured cost); lopsided ones to
    PHAST over the smaller side; the cells are the same."""
    g = grid_small
    rng = np.random.default_rng(9)
    r = router.MatrixRouter(MatStub(g))
    o = Oracle(g)
    for S, T in ((200, 150), (3, 400), (200, 150)):
        s = rng.integers(0, g.n, S).astype(np.uint32)
        t = rng.integers(0, g.n, T).astype(np.uint32)
        got = r.compute_matrix_block_gated(s, t, 1500)
        exp = o.compute_matrix_flat(s, t)
        assert np.array_equal(got, np.where(exp > 1500, INF, exp))
        if (S, T) == (3, 400):
            assert r.last_side == "forward"
    assert r.up_ns.load() > 0 and r.cell_ns.load() > 0
    # with both measured, the estimate decides: a cheap join keeps squares on the bucket
    r.fwd_ns.value, r.up_ns.value, r.cell_ns.value = 10**6, 10**3, 1
    assert r.side(1000, 1000) == "bucket" and r.side(2, 10**6) == "forward"
    assert router.MatrixRouter(MatStub(g), bucket=False).side(1000, 1000) == "forward"


@pytest.mark.gpu
def test_matrix_router_on_gpu(grid_config1):
    from bfgpu import BatchedPhastEngine
    g = grid_config1
    rng = np.random.default_rng(4)
    o = Oracle(g)
    with BatchedPhastEngine(g) as e:
        r = router.MatrixRouter(e)
        for S, T in ((300, 20), (20, 300), (64, 64), (200, 300)):
            s = rng.integers(0, g.n, S).astype(np.uint32)
            t = rng.integers(0, g.n, T).astype(np.uint32)
            got = r.compute_matrix_block_gated(s, t)
            assert np.array_equal(got, o.compute_matrix_flat_mt(s, t)), (S, T, r.last_side)
