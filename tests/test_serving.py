"""The serving loop's host logic: admission, coalescing into shared sweeps and
the split back into per-request results.  CPU tests run it over an engine
whose sweeps are the oracle's; the GPU test runs it over the HIP engine and
checks every result against the same request issued alone."""
import time

import numpy as np
import pytest

from bfgpu.graph import INF
from bfgpu.serving import IsochroneRequest, MatrixRequest, PlainLoop, ServingLoop, run_plain, run_round
from oracle.oracle import Oracle


class OracleEngine:
    """The engine calls the serving loop makes, answered by the oracle."""

    def __init__(self, g):
        self.o, self.n_nodes = Oracle(g), g.n
        self.calls = 0

    def info(self):
        return {"n_nodes": self.n_nodes, "n_top": self.n_nodes // 4}

    def free_device_bytes(self):
        return 1 << 34

    def compute_matrix_block_gated(self, sources, targets, threshold):
        self.calls += 1
        m = self.o.compute_matrix_flat_mt(sources, targets)
        if threshold != INF:
            m = np.where(m > threshold, INF, m).astype(np.uint32)
        return m, None

    def compute_matrix_multi(self, requests):
        self.calls += 1
        out = []
        for s, t, thr in requests:
            m = self.o.compute_matrix_flat_mt(s, t)
            if thr != INF:
                m = np.where(m > thr, INF, m).astype(np.uint32)
            out.append(m)
        return out

    def isochrone_lists(self, seed_lists, thresholds):
        self.calls += 1
        return [np.array(self.o.seeded(seeds, int(thr)), dtype=np.uint32).reshape(-1, 2)
                for seeds, thr in zip(seed_lists, thresholds)]


def mixed_stream(g, rng, count):
    reqs = []
    for _ in range(count):
        if rng.random() < 0.7:
            s = rng.integers(0, g.n, int(rng.integers(1, 40))).astype(np.uint32)
            t = rng.integers(0, g.n, int(rng.integers(1, 50))).astype(np.uint32)
            thr = INF if rng.random() < 0.7 else 1500
            reqs.append(MatrixRequest(s, t, thr))
        else:
            reqs.append(IsochroneRequest(int(rng.integers(0, g.n)), int(rng.choice([300, 600, 1800]))))
    return reqs


def alone(engine, reqs):
    return [run_round(engine, [r])[0] for r in reqs]


def test_round_results_equal_requests_alone(grid_small):
    eng = OracleEngine(grid_small)
    reqs = mixed_stream(grid_small, np.random.default_rng(1), 16)
    together = run_round(eng, reqs)
    assert eng.calls <= 2  # one matrix sweep, one isochrone sweep
    for a, b in zip(together, alone(eng, reqs)):
        assert np.array_equal(a, b)


def test_matrix_request_against_the_oracle(grid_small):
    eng = OracleEngine(grid_small)
    rng = np.random.default_rng(2)
    r = MatrixRequest(rng.integers(0, grid_small.n, 9).astype(np.uint32),
                      np.array([5, 3, 5, 0], dtype=np.uint32))  # duplicate target columns
    got = run_round(eng, [r, MatrixRequest(r.targets, r.sources)])[0]
    assert np.array_equal(got, eng.o.compute_matrix_flat(r.sources, r.targets))


@pytest.mark.parametrize("loop_cls", [ServingLoop, PlainLoop])
def test_loops_serve_a_burst(grid_small, loop_cls):
    eng = OracleEngine(grid_small)
    reqs = mixed_stream(grid_small, np.random.default_rng(3), 24)
    want = alone(OracleEngine(grid_small), reqs)
    with loop_cls(eng) as loop:
        futs = [loop.submit(r) for r in reqs]
        got = [f.result(timeout=120) for f in futs]
        rounds = loop.rounds
    for a, b in zip(got, want):
        assert np.array_equal(a, b)
    if loop_cls is ServingLoop:
        assert rounds < len(reqs)  # requests that waited were coalesced
    else:
        assert rounds == len(reqs)


def test_admission_limits_a_round(grid_small):
    """A round takes the oldest requests while their lanes fit the budget."""
    from concurrent.futures import Future
    eng = OracleEngine(grid_small)
    reqs = [MatrixRequest(np.arange(10, dtype=np.uint32), np.arange(10, dtype=np.uint32)) for _ in range(6)]
    with ServingLoop(eng, slots=1, max_lanes=25) as loop:
        with loop._cv:  # the worker needs this lock to take requests: queue all six first
            pending = [(r, Future()) for r in reqs]
            loop._pending.extend(pending)
            batch, res = loop._take()
            loop._release(res)
            loop._pending.extendleft(reversed(batch))
            loop._cv.notify()
        assert len(batch) == 2  # 10 + 10 lanes fit 25, a third request would not
        for _, f in pending:
            f.result(timeout=60)


@pytest.mark.gpu
def test_serving_on_gpu_equals_requests_alone(grid_config1):
    from bfgpu import BatchedPhastEngine
    g = grid_config1
    reqs = mixed_stream(g, np.random.default_rng(4), 30)
    print("[gpu-test] serving burst", len(reqs), flush=True)
    with BatchedPhastEngine(g) as e:
        want = [run_plain(e, r) for r in reqs]
        with ServingLoop(e) as loop:
            got = [f.result(timeout=120) for f in [loop.submit(r) for r in reqs]]
            rounds = loop.rounds
    assert rounds < len(reqs)
    for a, b in zip(got, want):
        assert np.array_equal(a, b)
    o = Oracle(g)
    for r, res in zip(reqs, got):
        if isinstance(r, MatrixRequest):
            m = o.compute_matrix_flat_mt(r.sources, r.targets)
            if r.threshold != INF:
                m = np.where(m > r.threshold, INF, m)
            assert np.array_equal(res, m)
        else:
            assert res.tolist() == [list(x) for x in o.seeded([(r.origin, 0)], r.limit)]


class FailingEngine(OracleEngine):
    """An engine whose calls fail while `fail` is set, the way a device call
    fails (BFG_OUT_OF_MEMORY forced from the host)."""

    def __init__(self, g):
        super().__init__(g)
        self.fail = False

    def _maybe_fail(self):
        if self.fail:
            from bfgpu.engine import BfgError
            raise BfgError(2, "forced out of memory")

    def compute_matrix_block_gated(self, sources, targets, threshold):
        self._maybe_fail()
        return super().compute_matrix_block_gated(sources, targets, threshold)

    def compute_matrix_multi(self, requests):
        self._maybe_fail()
        return super().compute_matrix_multi(requests)

    def isochrone_lists(self, seed_lists, thresholds):
        self._maybe_fail()
        return super().isochrone_lists(seed_lists, thresholds)


@pytest.mark.parametrize("loop_cls", [ServingLoop, PlainLoop])
def test_cancelled_request_leaves_the_loop_running(grid_small, loop_cls):
    """A request cancelled by its client before its round never reaches the
    engine, and the loop keeps serving the requests after it."""
    eng = OracleEngine(grid_small)
    reqs = mixed_stream(grid_small, np.random.default_rng(5), 6)
    want = alone(OracleEngine(grid_small), reqs)
    with loop_cls(eng) as loop:
        with loop._cv:  # hold the worker off while the requests queue up
            futs = [loop.submit(r) for r in reqs]
            assert futs[2].cancel()
        got = [f.result(timeout=60) if k != 2 else None for k, f in enumerate(futs)]
        assert futs[2].cancelled()
        late = loop.submit(reqs[0]).result(timeout=60)
    for k, (a, b) in enumerate(zip(got, want)):
        if k != 2:
            assert np.array_equal(a, b)
    assert np.array_equal(late, want[0])


@pytest.mark.parametrize("loop_cls", [ServingLoop, PlainLoop])
def test_failed_round_fails_only_its_requests(grid_small, loop_cls):
    """A round whose device call fails (out of memory forced from the host)
    fails the requests of that round with the engine's error; requests
    submitted afterwards are served normally."""
    from bfgpu.engine import BfgError
    eng = FailingEngine(grid_small)
    reqs = mixed_stream(grid_small, np.random.default_rng(6), 4)
    want = alone(OracleEngine(grid_small), reqs)
    with loop_cls(eng) as loop:
        eng.fail = True
        bad = loop.submit(reqs[0])
        with pytest.raises(BfgError):
            bad.result(timeout=60)
        eng.fail = False
        got = [f.result(timeout=60) for f in [loop.submit(r) for r in reqs]]
    for a, b in zip(got, want):
        assert np.array_equal(a, b)


class SlowEngine(OracleEngine):
    """Records when each engine call starts and ends; each call takes `delay`
    seconds (a round on the device)."""

    def __init__(self, g, delay=0.3):
        super().__init__(g)
        self.delay = delay
        self.spans = []

    def _timed(self, fn, *a):
        t0 = time.perf_counter()
        time.sleep(self.delay)
        r = fn(*a)
        self.spans.append((t0, time.perf_counter()))
        return r

    def compute_matrix_multi(self, requests):
        return self._timed(super().compute_matrix_multi, requests)

    def isochrone_lists(self, seed_lists, thresholds):
        return self._timed(super().isochrone_lists, seed_lists, thresholds)


def test_request_arriving_mid_round_is_admitted_and_overlaps(grid_small):
    """Continuous batching: a request that arrives while a round runs starts
    in a free slot at once, and the two rounds overlap."""
    eng = SlowEngine(grid_small)
    rng = np.random.default_rng(7)
    a = MatrixRequest(rng.integers(0, grid_small.n, 5).astype(np.uint32), rng.integers(0, grid_small.n, 7).astype(np.uint32))
    b = MatrixRequest(rng.integers(0, grid_small.n, 3).astype(np.uint32), rng.integers(0, grid_small.n, 4).astype(np.uint32))
    with ServingLoop(eng, slots=2) as loop:
        fa = loop.submit(a)
        time.sleep(0.1)  # a's round is on the "device"
        fb = loop.submit(b)
        ra, rb = fa.result(timeout=30), fb.result(timeout=30)
    (a0, a1), (b0, b1) = sorted(eng.spans)
    assert b0 < a1  # b started before a's round ended
    assert np.array_equal(ra, eng.o.compute_matrix_flat(a.sources, a.targets))
    assert np.array_equal(rb, eng.o.compute_matrix_flat(b.sources, b.targets))


def test_memory_budget_holds_a_request_until_a_round_frees(grid_small):
    """Admission by device memory: a request whose bytes do not fit what the
    round in flight leaves of the budget waits for that round, even with a
    free slot."""
    from bfgpu.serving import request_bytes
    eng = SlowEngine(grid_small)
    rng = np.random.default_rng(8)
    a = MatrixRequest(rng.integers(0, grid_small.n, 20).astype(np.uint32), rng.integers(0, grid_small.n, 20).astype(np.uint32))
    b = MatrixRequest(rng.integers(0, grid_small.n, 20).astype(np.uint32), rng.integers(0, grid_small.n, 20).astype(np.uint32))
    info = eng.info()
    need = request_bytes(a, info["n_nodes"], info["n_top"])
    with ServingLoop(eng, slots=2, mem_budget=int(1.5 * need)) as loop:
        fa = loop.submit(a)
        time.sleep(0.1)
        fb = loop.submit(b)
        fa.result(timeout=30)
        fb.result(timeout=30)
    (a0, a1), (b0, b1) = sorted(eng.spans)
    assert b0 >= a1  # b waited for a's memory


class BucketOracleEngine(OracleEngine):
    """OracleEngine with the bucket side (the oracle's bucket many-to-many)."""

    def __init__(self, g):
        super().__init__(g)
        self.bucket_calls = 0

    def table_bucket_parallel_bounded(self, sources, targets, threshold=INF):
        self.bucket_calls += 1
        m = self.o.table_bucket(sources, targets, threshold)
        return np.where(m > threshold, INF, m).astype(np.uint32), None


def test_round_routes_large_matrices_to_the_bucket(grid_small):
    """A round sends matrices with at least bucket_min sources and targets to
    the bucket and the rest to the shared PHAST sweep; results equal the
    requests alone either way; PlainLoop routes through a MatrixRouter."""
    g = grid_small
    rng = np.random.default_rng(17)
    eng = BucketOracleEngine(g)
    big = MatrixRequest(rng.integers(0, g.n, 12).astype(np.uint32), rng.integers(0, g.n, 15).astype(np.uint32), 1500)
    small = MatrixRequest(rng.integers(0, g.n, 3).astype(np.uint32), rng.integers(0, g.n, 40).astype(np.uint32))
    iso = IsochroneRequest(int(rng.integers(0, g.n)), 600)
    out = run_round(eng, [big, small, iso], bucket_min=10)
    assert eng.bucket_calls == 1
    for r, got in zip([big, small, iso], out):
        assert np.array_equal(got, run_plain(OracleEngine(g), r))
    with PlainLoop(eng) as loop:
        assert loop.router is not None
        got = loop.submit(big).result(timeout=60)
    assert np.array_equal(got, out[0])
