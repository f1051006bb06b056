"""Shape routing of matrix calls.

`phast_wins` restates the reference's router predicate (route/src/matrix/
bucket_ch.rs:6677-6711) with its measured cost cells (SWEEP_NS, SCAN_NS,
:6660-6670, EWMA alpha 1/4) and its BUTTERFLY_MATRIX_ALGO override:
- "bucket" -> never PHAST; "phast" -> PHAST iff S <= T;
- S > T -> not PHAST (the reverse-field variant is the reference's follow-up);
- with both cells measured: PHAST iff S * scan < (S + T) * sweep;
- otherwise the structural fallback: scan = n, sweep = clamp(n / 400, 500,
  20 000).
`table_seeded_bounded_routed` routes a seeded table the way the reference
does (:6535-6574): the PHAST path is the GPU table, the bucket path the
caller's bucket engine when one is given (it is not part of this build; the
GPU table serves both when none is given).

`MatrixRouter` is the GPU's own routing of plain matrices.  The same cells
come from three engines: PHAST lanes over the sources on the forward CCH,
PHAST lanes over the targets on the transposed CCH (all-to-one rows,
range/phast.rs:869-885), or the bucket many-to-many (table_bucket_parallel,
bucket_ch.rs:3436-3520: upward phases of both sides joined at their common
ancestors).  It keeps EWMA cost cells measured on the calls it routes: ns per
lane of each PHAST side, and for the bucket ns per upward lane (S + T) plus ns
per cell of the join (S * T), and sends each call to the cheapest estimate.
Before a side is measured: the bucket when both sides have at least
BUCKET_MIN lanes, otherwise PHAST over the side with fewer lanes.
"""
import os
import threading
import time

import numpy as np

from .graph import INF


class CostCell:
    """A cost cell in nanoseconds (0 = not measured yet)."""

    def __init__(self):
        self.value = 0
        self._lock = threading.Lock()

    def load(self):
        return self.value


def update_cost_ewma(cell, sample_ns):
    """bucket_ch.rs:6663-6671: the first sample is taken as is, then
    prev - prev/4 + sample/4 (integer division), never below 1."""
    with cell._lock:
        prev = cell.value
        nxt = sample_ns if prev == 0 else prev - prev // 4 + sample_ns // 4
        cell.value = max(int(nxt), 1)


SWEEP_NS = CostCell()  # per bucket sweep (the caller's bucket engine)
SCAN_NS = CostCell()   # per PHAST scan (one source's field)


def phast_wins(n_sources, n_targets, n_nodes, sweep_cell=SWEEP_NS, scan_cell=SCAN_NS):
    """bucket_ch.rs:6677-6711: does min(S, T) full scans beat (S + T) sweeps?"""
    algo = os.environ.get("BUTTERFLY_MATRIX_ALGO")
    if algo == "bucket":
        return False
    if algo == "phast":
        return n_sources <= n_targets
    if n_sources > n_targets:
        return False  # the reverse-field variant is a follow-up in the reference
    sweep, scan = sweep_cell.load(), scan_cell.load()
    if sweep > 0 and scan > 0:
        scan_cost, sweep_cost = scan, sweep
    else:
        scan_cost, sweep_cost = n_nodes, min(max(n_nodes // 400, 500), 20_000)
    return n_sources * scan_cost < (n_sources + n_targets) * sweep_cost


def table_seeded_bounded_routed(engine, src_seedsets, tgt_seedsets, threshold=INF, bucket=None):
    """bucket_ch.rs:6535-6574 over the GPU engine: an empty side gives an
    all-INF table; phast_wins -> the GPU PHAST table (table_phast_lopsided
    semantics), whose per-source scan time feeds SCAN_NS (:6769); otherwise
    bucket(src_seedsets, tgt_seedsets, threshold) when given (its caller
    feeds SWEEP_NS), else the GPU table as well.  Returns (S*T u32 cells,
    S*T u8 same-edge conflicts) like engine.table_seeded_bounded_routed."""
    S, T = len(src_seedsets), len(tgt_seedsets)
    if S == 0 or T == 0:
        return np.full(S * T, INF, dtype=np.uint32), np.zeros(S * T, dtype=np.uint8)
    if bucket is not None and not phast_wins(S, T, engine.n_nodes):
        return bucket(src_seedsets, tgt_seedsets, threshold)
    t0 = time.perf_counter_ns()
    res = engine.table_seeded_bounded_routed(src_seedsets, tgt_seedsets, threshold)
    update_cost_ewma(SCAN_NS, (time.perf_counter_ns() - t0) // S)
    return res


BUCKET_MIN = 128  # lanes on each side from which the bucket is tried before it is measured


class MatrixRouter:
    """Routes plain matrices to forward PHAST (source lanes), reverse PHAST
    (target lanes, transposed CCH) or the bucket many-to-many by measured
    cost."""

    def __init__(self, engine, bucket=True):
        self.engine = engine
        self.bucket = bucket
        self.fwd_ns, self.rev_ns = CostCell(), CostCell()  # ns per lane of each PHAST side
        self.up_ns, self.cell_ns = CostCell(), CostCell()  # bucket: ns per upward lane, ns per joined cell
        self.last_side = None

    def estimates(self, n_sources, n_targets):
        """Estimated ns of each side (None where not measured yet)."""
        f, r = self.fwd_ns.load(), self.rev_ns.load()
        u, c = self.up_ns.load(), self.cell_ns.load()
        return {"forward": n_sources * f if f else None,
                "reverse": n_targets * r if r else None,
                "bucket": ((n_sources + n_targets) * u + n_sources * n_targets * c) if (u and c) else None}

    def side(self, n_sources, n_targets):
        S, T = n_sources, n_targets
        f, r = self.fwd_ns.load(), self.rev_ns.load()
        f, r = (f or r), (r or f)  # one measured PHAST side stands in for the other
        u, c = self.up_ns.load(), self.cell_ns.load()
        have_b = self.bucket and bool(u and c)
        if f:
            cost = {"forward": S * f, "reverse": T * r}
            phast = min(cost, key=cost.get)  # forward on a tie
        else:
            phast = "forward" if S <= T else "reverse"
        if f and have_b:
            return "bucket" if (S + T) * u + S * T * c < cost[phast] else phast
        if self.bucket and not have_b and min(S, T) >= BUCKET_MIN:
            return "bucket"
        return phast

    def compute_matrix_block_gated(self, sources, targets, threshold=INF):
        """The S x T matrix (row-major u32), equal to the engine's own
        compute_matrix_block_gated, through the cheapest side."""
        src = np.ascontiguousarray(sources, dtype=np.uint32)
        tgt = np.ascontiguousarray(targets, dtype=np.uint32)
        side = self.side(src.size, tgt.size)
        t0 = time.perf_counter_ns()
        if side == "forward":
            m, _ = self.engine.compute_matrix_block_gated(src, tgt, threshold)
            lanes, cell = src.size, self.fwd_ns
        elif side == "reverse":
            rev = self.engine.reverse_engine()
            mt, _ = rev.compute_matrix_block_gated(tgt, src, threshold)
            m = np.ascontiguousarray(mt.reshape(tgt.size, src.size).T).reshape(-1)
            lanes, cell = tgt.size, self.rev_ns
        else:
            m, st = self.engine.table_bucket_parallel_bounded(src, tgt, threshold)
            lanes, cell = 0, None
            total = time.perf_counter_ns() - t0
            if src.size and tgt.size:
                # the join's share of the call by the engine's own phase times
                up, join = float(st.upward_time_ms), float(st.downward_time_ms)
                share = join / (up + join) if up + join > 0 else 0.5
                update_cost_ewma(self.up_ns, int(total * (1 - share)) // (src.size + tgt.size))
                update_cost_ewma(self.cell_ns, max(int(total * share) // (src.size * tgt.size), 1))
        if lanes:
            update_cost_ewma(cell, (time.perf_counter_ns() - t0) // lanes)
        self.last_side = side
        return m
