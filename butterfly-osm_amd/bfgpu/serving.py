"""Serving loop over one device-resident engine: continuous batching of matrix
and isochrone requests.

The graph and the label arenas stay on the GPU for the life of the engine.
Requests arrive on any thread (submit returns a Future).  The loop has
`slots` workers; each runs one round at a time, and the engine runs each
worker's calls on a stream and label arena of their own, so up to `slots`
rounds overlap on the device.  Whenever a slot is free and requests wait, it
admits the oldest pending requests while their lanes fit `max_lanes` and
their estimated device memory fits what the rounds in flight leave of the
budget (a fraction of the device's free memory when the loop starts), and
runs them as one round:
- every matrix request with fewer than `bucket_min` sources or targets
  becomes lanes of one shared PHAST sweep (bfg_matrix_multi), each request
  gathering only its own rows and columns; larger ones run through the bucket
  many-to-many (bfg_table_bucket: upward phases of both sides and a join),
  whose cost grows with S + T and S * T rather than with a sweep per source
  group, when the engine has it;
- every isochrone becomes one lane of one block-gated sweep with its own
  limit (bfg_isochrone_lists), answered with its settled (rank, time) list.
Requests that arrive while rounds are on the GPU are admitted as soon as a
slot and memory free up.  A cell or a settled node does not depend on what
else shares its sweep (labels are exact distances), so every result is
bit-identical to the same request issued alone.

PlainLoop serves the same stream the plain way, one GPU call per request in
arrival order, each matrix routed by a MatrixRouter (forward PHAST, reverse
PHAST or the bucket, by measured cost); the serving benchmark (bench.py
--serving) compares the two.
"""
import collections
import dataclasses
import threading
import time
from concurrent.futures import Future

import numpy as np

from .graph import INF
from .router import MatrixRouter


@dataclasses.dataclass
class MatrixRequest:
    """An S x T duration matrix (compute_matrix_block_gated semantics)."""
    sources: np.ndarray
    targets: np.ndarray
    threshold: int = INF

    def lanes(self):
        return len(self.sources)


@dataclasses.dataclass
class IsochroneRequest:
    """One isochrone: the settled (rank, time <= limit) pairs in ascending rank
    (run_phast_bounded_fast_seeded's output)."""
    origin: int
    limit: int

    def lanes(self):
        return 1


BUCKET_MIN = 128  # sources and targets from which a round's matrix runs through the bucket


def run_round(engine, reqs, bucket_min=BUCKET_MIN):
    """Results of one round of requests, in the order given: a row-major S x T
    u32 array per matrix request, an (m, 2) u32 array of (rank, time) per
    isochrone.  Matrices with at least bucket_min sources and targets run
    through the engine's bucket many-to-many when it has one (0 = never)."""
    out = [None] * len(reqs)
    mats = [i for i, r in enumerate(reqs) if isinstance(r, MatrixRequest)]
    isos = [i for i, r in enumerate(reqs) if not isinstance(r, MatrixRequest)]
    if bucket_min and hasattr(engine, "table_bucket_parallel_bounded"):
        big = [i for i in mats if min(len(reqs[i].sources), len(reqs[i].targets)) >= bucket_min]
        for i in big:
            out[i] = engine.table_bucket_parallel_bounded(reqs[i].sources, reqs[i].targets, reqs[i].threshold)[0]
        mats = [i for i in mats if out[i] is None]
    if mats:
        res = engine.compute_matrix_multi([(reqs[i].sources, reqs[i].targets, reqs[i].threshold) for i in mats])
        for i, m in zip(mats, res):
            out[i] = m
    if isos:
        lists = engine.isochrone_lists([[(int(reqs[i].origin), 0)] for i in isos],
                                       np.array([reqs[i].limit for i in isos], dtype=np.uint32))
        for i, lst in zip(isos, lists):
            out[i] = lst
    return out


def run_plain(engine, req, router=None):
    """One request as its own plain engine call (the per-request baseline); a
    matrix goes through `router` (a MatrixRouter) when given."""
    if isinstance(req, MatrixRequest):
        if router is not None:
            return router.compute_matrix_block_gated(req.sources, req.targets, req.threshold)
        return engine.compute_matrix_block_gated(req.sources, req.targets, req.threshold)[0]
    return engine.isochrone_lists([[(int(req.origin), 0)]], np.array([req.limit], dtype=np.uint32))[0]


# Estimated settled fraction of an isochrone's nodes (sizes its list buffer).
_ISO_SETTLED = 0.25


def request_bytes(req, n_nodes, n_top, bucket_min=0):
    """Device bytes a request adds to a round: its label lanes (a PHAST
    matrix's lanes hold the top region plus the bottom paths above its
    targets, about 24 nodes per target; a bucket matrix's source and target
    lanes hold the top region; an isochrone's lane holds every node) and its
    outputs."""
    if isinstance(req, MatrixRequest):
        S, T = len(req.sources), len(req.targets)
        if bucket_min and min(S, T) >= bucket_min:
            return 4 * (S + T) * n_top + 4 * S * T
        stride = min(n_nodes, n_top + 24 * T)
        return 4 * S * stride + 4 * S * T
    words = (n_nodes + 63) // 64
    return 4 * n_nodes + 8 * words + int(8 * _ISO_SETTLED * n_nodes)


class _Loop:
    """Shared machinery: a pending queue, `slots` worker threads, futures
    that fail (not hang) when a round fails, and cancelled requests dropped."""

    def __init__(self, engine, slots):
        self.engine = engine
        self._pending = collections.deque()
        self._cv = threading.Condition()
        self._stop = False
        self.rounds = 0
        self._workers = [threading.Thread(target=self._run, daemon=True) for _ in range(slots)]
        for t in self._workers:
            t.start()

    def submit(self, req):
        fut = Future()
        fut.submitted = time.perf_counter()
        with self._cv:
            self._pending.append((req, fut))
            self._cv.notify()
        return fut

    def close(self):
        with self._cv:
            self._stop = True
            self._cv.notify_all()
        for t in self._workers:
            t.join()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _take(self):
        """Called with the lock held: the next round's (request, future) pairs
        and its reservation, or (None, 0) when nothing can be admitted now."""
        raise NotImplementedError

    def _release(self, reservation):
        pass

    def _serve(self, reqs):
        return run_round(self.engine, reqs)

    def _run(self):
        while True:
            with self._cv:
                while True:
                    if not self._pending:
                        if self._stop:
                            return
                        self._cv.wait()
                        continue
                    batch, reservation = self._take()
                    if batch:
                        break
                    self._cv.wait()  # a round in flight holds what this one needs
            # a request its client cancelled before the round starts leaves
            # the round; the rest can no longer be cancelled
            batch = [(r, f) for r, f in batch if f.set_running_or_notify_cancel()]
            try:
                if batch:
                    try:
                        res = self._serve([r for r, _ in batch])
                    except BaseException as e:  # a failed round fails its requests, the loop goes on
                        for _, f in batch:
                            if not f.done():
                                f.set_exception(e)
                        continue
                    done = time.perf_counter()
                    with self._cv:
                        self.rounds += 1
                    for (_, f), r in zip(batch, res):
                        f.completed = done
                        if not f.done():
                            f.set_result(r)
            finally:
                with self._cv:
                    self._release(reservation)
                    self._cv.notify_all()


class ServingLoop(_Loop):
    """Continuous batching: `slots` rounds may be on the device at once; a
    round takes the oldest pending requests while their lanes fit max_lanes
    and their estimated device bytes fit the budget left by the rounds in
    flight (with no round in flight the oldest request is always taken).
    The budget is mem_fraction of the device's free memory when the loop
    starts, or mem_budget bytes when given."""

    def __init__(self, engine, slots=3, max_lanes=8192, mem_fraction=0.5, mem_budget=None, bucket_min=BUCKET_MIN):
        self.max_lanes = max_lanes
        self.bucket_min = bucket_min
        self._bucket_min = bucket_min if hasattr(engine, "table_bucket_parallel_bounded") else 0
        if hasattr(engine, "set_concurrency"):
            engine.set_concurrency(slots)
        info = engine.info() if hasattr(engine, "info") else {}
        self._n = int(info.get("n_nodes", getattr(engine, "n_nodes", 0)))
        self._n_top = int(info.get("n_top", self._n))
        if mem_budget is None:
            free = engine.free_device_bytes() if hasattr(engine, "free_device_bytes") else 8 << 30
            mem_budget = int(mem_fraction * free)
        self.mem_budget = mem_budget
        self._inflight_bytes = 0
        self._inflight = 0
        super().__init__(engine, slots)

    def _take(self):
        batch, lanes, nbytes = [], 0, 0
        for req, fut in self._pending:
            l, b = req.lanes(), request_bytes(req, self._n, self._n_top, self._bucket_min)
            fits = lanes + l <= self.max_lanes and self._inflight_bytes + nbytes + b <= self.mem_budget
            if not fits and (batch or self._inflight):
                break
            batch.append((req, fut))
            lanes += l
            nbytes += b
        if not batch:
            return None, 0
        for _ in batch:
            self._pending.popleft()
        self._inflight_bytes += nbytes
        self._inflight += 1
        return batch, nbytes

    def _release(self, reservation):
        if reservation is None:
            return
        self._inflight_bytes -= reservation
        self._inflight -= 1

    def _serve(self, reqs):
        return run_round(self.engine, reqs, self.bucket_min)


class PlainLoop(_Loop):
    """One GPU call per request, in arrival order (one slot); matrices routed
    by a MatrixRouter when the engine has the bucket side (route=True)."""

    def __init__(self, engine, route=True):
        self.router = MatrixRouter(engine) if route and hasattr(engine, "table_bucket_parallel_bounded") else None
        super().__init__(engine, 1)

    def _take(self):
        return [self._pending.popleft()], None

    def _serve(self, reqs):
        return [run_plain(self.engine, r, self.router) for r in reqs]
