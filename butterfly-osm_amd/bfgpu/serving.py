"""Serving loop over one device-resident engine: continuous batching of matrix
and isochrone requests.

The graph and the label arenas stay on the GPU for the life of the engine.
Requests arrive on any thread (submit returns a Future).  One worker thread
runs rounds; a round admits the oldest pending requests while they fit a lane
budget and a device-memory budget, coalesces them and runs them as shared
sweeps:
- every admitted matrix request's sources become lanes of one matrix sweep
  (per threshold) over the union of their targets, and each request reads
  back its own rows and columns;
- every admitted isochrone becomes one lane of one isochrone sweep with its
  own limit.
Requests that arrive while a round is on the GPU are admitted by the next
round, which starts as soon as the device is free.  A cell or a reach bit
does not depend on what else shares its sweep (labels are exact distances),
so every result is bit-identical to the same request issued alone.

PlainLoop serves the same stream the plain way, one GPU call per request in
arrival order; the serving benchmark (bench.py --serving) compares the two.
"""
import collections
import dataclasses
import threading
import time
from concurrent.futures import Future

import numpy as np

from .graph import INF


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
    """One block-gated isochrone: reach bits of the nodes within limit."""
    origin: int
    limit: int

    def lanes(self):
        return 1


def run_round(engine, reqs):
    """Results of one round of requests, in the order given: a row-major S x T
    u32 array per matrix request, a ceil(n/64) u64 bit row per isochrone."""
    out = [None] * len(reqs)
    by_thr = collections.defaultdict(list)
    isos = []
    for i, r in enumerate(reqs):
        if isinstance(r, MatrixRequest):
            by_thr[int(r.threshold)].append(i)
        else:
            isos.append(i)
    for thr, idx in by_thr.items():
        srcs = [np.asarray(reqs[i].sources, dtype=np.uint32) for i in idx]
        tgts = [np.asarray(reqs[i].targets, dtype=np.uint32) for i in idx]
        S = np.concatenate(srcs)
        T, inv = np.unique(np.concatenate(tgts), return_inverse=True)
        m, _ = engine.compute_matrix_block_gated(S, T, thr)
        m = m.reshape(S.size, T.size)
        r0 = c0 = 0
        for i, s, t in zip(idx, srcs, tgts):
            out[i] = m[r0:r0 + s.size][:, inv[c0:c0 + t.size]].reshape(-1).copy()
            r0 += s.size
            c0 += t.size
    if isos:
        org = np.array([reqs[i].origin for i in isos], dtype=np.uint32)
        lim = np.array([reqs[i].limit for i in isos], dtype=np.uint32)
        bits, _ = engine.isochrones(org, lim, want_times=False)
        for k, i in enumerate(isos):
            out[i] = bits[k].copy()
    return out


def request_bytes(req, n_nodes):
    """Device output bytes a request adds to a round (the admission measure)."""
    if isinstance(req, MatrixRequest):
        return 4 * len(req.sources) * len(req.targets)
    return 8 * ((n_nodes + 63) // 64)


class _Loop:
    def __init__(self, engine):
        self.engine = engine
        self._pending = collections.deque()
        self._cv = threading.Condition()
        self._stop = False
        self.rounds = 0
        self._worker = threading.Thread(target=self._run, daemon=True)
        self._worker.start()

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
            self._cv.notify()
        self._worker.join()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _take(self):
        raise NotImplementedError

    def _run(self):
        while True:
            with self._cv:
                while not self._pending and not self._stop:
                    self._cv.wait()
                if not self._pending:
                    return
                batch = self._take()
            # a request its client cancelled before the round starts leaves the
            # round; the rest can no longer be cancelled
            batch = [(r, f) for r, f in batch if f.set_running_or_notify_cancel()]
            if not batch:
                continue
            try:
                res = run_round(self.engine, [r for r, _ in batch])
            except BaseException as e:  # a failed round fails its requests, the loop goes on
                for _, f in batch:
                    if not f.done():
                        f.set_exception(e)
                continue
            self.rounds += 1
            done = time.perf_counter()
            for (_, f), r in zip(batch, res):
                f.completed = done
                if not f.done():
                    f.set_result(r)


class ServingLoop(_Loop):
    """Continuous batching: each round takes the oldest pending requests while
    their lanes fit max_lanes and their outputs fit max_bytes (the oldest
    request is always taken)."""

    def __init__(self, engine, max_lanes=8192, max_bytes=4 << 30):
        self.max_lanes, self.max_bytes = max_lanes, max_bytes
        super().__init__(engine)

    def _take(self):
        batch, lanes, nbytes = [], 0, 0
        while self._pending:
            req, fut = self._pending[0]
            l, b = req.lanes(), request_bytes(req, self.engine.n_nodes)
            if batch and (lanes + l > self.max_lanes or nbytes + b > self.max_bytes):
                break
            batch.append(self._pending.popleft())
            lanes += l
            nbytes += b
        return batch


class PlainLoop(_Loop):
    """One GPU call per request, in arrival order."""

    def _take(self):
        return [self._pending.popleft()]
