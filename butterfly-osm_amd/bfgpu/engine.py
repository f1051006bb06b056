"""Host mirror of the reference's K-lane batched PHAST API over the C ABI.

`BatchedPhastEngine` keeps the method names, argument meaning and error
behaviour of route/src/matrix/batched_phast.rs (sources and targets are CCH
ranks, matrices are row-major u32, unreachable = u32::MAX), with the work done
by the HIP engine in libbfgpu.so.  `run_phast_bounded_fast_seeded` mirrors
route/src/server/isochrone_handler.rs:338-491.
"""
import ctypes as C
import dataclasses

import numpy as np

from . import _native
from .graph import INF, CchGraph

K_LANES = 8  # the reference's lane count (batched_phast.rs:33); the GPU runs wider groups


BFG_BUFFER_TOO_SMALL = 6  # bfg_result: the output buffer holds less than the result


class BfgError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"bfgpu error {code}: {msg}")
        self.code = code


def _check(rc):
    if rc != 0:
        raise BfgError(rc, _native.bfgpu().bfg_last_error().decode())


@dataclasses.dataclass
class BatchedPhastStats:
    """batched_phast.rs:67-80, with device phase times."""
    n_sources: int = 0
    upward_time_ms: float = 0.0
    downward_time_ms: float = 0.0
    output_time_ms: float = 0.0
    total_time_ms: float = 0.0
    batches: int = 0
    sweep_launches: int = 0
    nodes_swept: int = 0  # nodes / DOWN arcs one lane group's sweep covered
    arcs_swept: int = 0
    dom_ms: float = 0.0  # the DOWN row kernel: summed launch time, launches, algorithmic bytes
    dom_launches: int = 0
    dom_bytes: int = 0
    gate_seg_units: int = 0  # block gate: (segment, wave group) units, and those left open
    gate_seg_open: int = 0
    gate_bt_units: int = 0   # the same for (bottom tree, wave group)
    gate_bt_open: int = 0
    join_ms: float = 0.0     # bucket matrices: the join kernel's time and label positions visited per tile
    join_visits: int = 0
    join_label_bits: int = 0  # the width the join ran at (16 or 32) and the tiles redone at 32 bits
    join_tiles_u32: int = 0


@dataclasses.dataclass
class BatchedPhastResult:
    """batched_phast.rs:56-65: dist[lane][node]."""
    dist: np.ndarray
    n_lanes: int
    stats: BatchedPhastStats


def _u32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.uint32))


class BatchedPhastEngine:
    """BatchedPhastEngine::new(topo, weights) (batched_phast.rs:89) on one GPU."""

    def __init__(self, graph: CchGraph, device: int = 0, transposed: bool = False):
        """transposed: the context over graph's transposed CCH (the reverse
        metric), which the library transposes itself (bfg_ctx_create_reverse);
        `graph` then builds CchGraph.transpose only when asked for."""
        self._graph, self._graph_of = (None, graph) if transposed else (graph, None)
        self._lib = _native.bfgpu()
        h = C.c_void_p()
        d = graph.desc()
        create = self._lib.bfg_ctx_create_reverse if transposed else self._lib.bfg_ctx_create
        _check(create(C.byref(d), device, C.byref(h)))
        self._h = h
        self.n_nodes = graph.n
        self._device = device
        self._rev = None

    @property
    def graph(self):
        if self._graph is None:
            self._graph = self._graph_of.transpose()
        return self._graph

    def close(self):
        if getattr(self, "_paired", False) and getattr(self, "_h", None):
            self._lib.bfg_ctx_pair_reverse(self._h, None)
            self._paired = False
        if getattr(self, "_rev", None) is not None:
            self._rev.close()
            self._rev = None
        if getattr(self, "_h", None):
            self._lib.bfg_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- introspection ------------------------------------------------------
    def info(self):
        i = _native.CtxInfo()
        _check(self._lib.bfg_ctx_get_info(self._h, C.byref(i)))
        return {f: getattr(i, f) for f, _ in i._fields_}

    def last_stats(self, n_sources=0):
        t = _native.Timings()
        _check(self._lib.bfg_ctx_last_timings(self._h, C.byref(t)))
        return BatchedPhastStats(n_sources=n_sources, upward_time_ms=t.up_ms, downward_time_ms=t.sweep_ms,
                                 output_time_ms=t.output_ms, total_time_ms=t.total_ms, batches=t.batches,
                                 sweep_launches=t.sweep_launches, nodes_swept=t.nodes_swept,
                                 arcs_swept=t.arcs_swept, dom_ms=t.dom_ms, dom_launches=t.dom_launches,
                                 dom_bytes=t.dom_bytes, gate_seg_units=t.gate_seg_units,
                                 gate_seg_open=t.gate_seg_open, gate_bt_units=t.gate_bt_units,
                                 gate_bt_open=t.gate_bt_open, join_ms=t.join_ms, join_visits=t.join_visits,
                                 join_label_bits=t.join_label_bits, join_tiles_u32=t.join_tiles_u32)

    def configure(self, lanes=0, lane_capacity=0):
        _check(self._lib.bfg_ctx_configure(self._h, lanes, lane_capacity))

    def free_device_bytes(self):
        """Free memory of the engine's device (hipMemGetInfo)."""
        from . import hip
        hip.hip().hipSetDevice(self._device)
        return hip.mem_info()[0]

    def set_label_width(self, bits):
        """16 (packed 16-bit DOWN rows where exact, the default) or 32
        (bfg_ctx_set_label_width); results are identical."""
        _check(self._lib.bfg_ctx_set_label_width(self._h, bits))

    def set_concurrency(self, max_calls):
        """Calls that may run on the device at once, each on its own stream
        and label arena (bfg_ctx_set_concurrency)."""
        _check(self._lib.bfg_ctx_set_concurrency(self._h, max_calls))

    def tune(self, plan_rows=0, tri_scalar_segs=None):
        """DOWN unit size and the level width from which intra-segment blocks
        are scalar-loaded (bfg_ctx_tune); None / 0 keep the defaults."""
        _check(self._lib.bfg_ctx_tune(self._h, plan_rows, 0 if tri_scalar_segs is None else tri_scalar_segs + 1))

    def rank_to_filtered(self):
        return self.graph.rank_to_filtered

    # -- queries --------------------------------------------------------------
    def query_batch(self, sources):
        """batched_phast.rs:131-260 (any number of sources): dist[lane][node]."""
        return self.query_batch_bounded(sources, INF)

    query_batch_soa = query_batch      # batched_phast.rs:470-609: same result
    query_batch_blocked = query_batch  # batched_phast.rs:455-459

    def query_batch_bounded(self, sources, threshold):
        """batched_phast.rs:271-445: exact within threshold, INF beyond."""
        src = _u32(sources)
        out = np.empty((src.size, self.n_nodes), dtype=np.uint32)
        _check(self._lib.bfg_query_batch(self._h, _native.ptr(src), src.size, int(threshold) & INF,
                                         _native.ptr(out)))
        return BatchedPhastResult(dist=out, n_lanes=src.size, stats=self.last_stats(src.size))

    query_batch_block_gated = query_batch_bounded  # batched_phast.rs:946-1218

    def compute_matrix_flat(self, sources, targets):
        """batched_phast.rs:667-706 -> (row-major S*T u32, stats)."""
        return self.compute_matrix_block_gated(sources, targets, INF)

    compute_matrix_flat_soa = compute_matrix_flat      # :756-797
    compute_matrix_flat_blocked = compute_matrix_flat  # :710-751

    def compute_matrix_block_gated(self, sources, targets, threshold, out=None):
        """batched_phast.rs:1223-1265: cells > threshold are INF.  `out`
        (optional) is a caller-owned contiguous u32 array of S*T cells, e.g.
        the `a` of a hip.PinnedArray; it is filled and returned."""
        src, tgt = _u32(sources), _u32(targets)
        if out is None:
            out = np.empty(src.size * tgt.size, dtype=np.uint32)
        elif out.dtype != np.uint32 or out.size != src.size * tgt.size or not out.flags.c_contiguous:
            raise ValueError("out must be a contiguous u32 array of S*T cells")
        _check(self._lib.bfg_matrix_bounded(self._h, _native.ptr(src), src.size, _native.ptr(tgt), tgt.size,
                                            int(threshold) & INF, _native.ptr(out)))
        return out, self.last_stats(src.size)

    def compute_matrix_multi(self, requests):
        """Several (sources, targets, threshold) matrix requests in one lane
        batch (bfg_matrix_multi): a list of row-major S_r*T_r u32 arrays,
        each equal to compute_matrix_block_gated of the request alone."""
        srcs = [_u32(r[0]) for r in requests]
        tgts = [_u32(r[1]) for r in requests]
        thr = _u32([int(r[2]) & INF if len(r) > 2 else INF for r in requests])
        so = np.zeros(len(requests) + 1, dtype=np.uint32)
        to = np.zeros(len(requests) + 1, dtype=np.uint32)
        so[1:] = np.cumsum([a.size for a in srcs])
        to[1:] = np.cumsum([b.size for b in tgts])
        sizes = [a.size * b.size for a, b in zip(srcs, tgts)]
        out = np.empty(max(sum(sizes), 1), dtype=np.uint32)
        S = np.concatenate(srcs) if srcs else np.zeros(0, np.uint32)
        T = np.concatenate(tgts) if tgts else np.zeros(0, np.uint32)
        P = _native.ptr
        _check(self._lib.bfg_matrix_multi(self._h, len(requests), P(so), P(_u32(S)), P(to), P(_u32(T)), P(thr),
                                          P(out)))
        res, at = [], 0
        for k in sizes:
            res.append(out[at:at + k].copy())
            at += k
        return res

    def compute_matrix_device(self, d_sources, n_src, d_targets, n_tgt, d_out, threshold=INF):
        """Device-resident matrix (bfg_matrix_device): raw HIP pointers (ints or c_void_p)."""
        _check(self._lib.bfg_matrix_device(self._h, d_sources, n_src, d_targets, n_tgt, int(threshold) & INF,
                                           d_out))
        return self.last_stats(n_src)

    def compute_matrix(self, sources, targets, callback):
        """batched_phast.rs:619-661: callback(first_source_idx, rows, chunk) per 8-source chunk."""
        src = _u32(sources)
        tgt = _u32(targets) if len(targets) else np.arange(self.n_nodes, dtype=np.uint32)
        mat, stats = self.compute_matrix_flat(src, tgt)
        mat = mat.reshape(src.size, tgt.size)
        for b in range(0, src.size, K_LANES):
            callback(b, list(mat[b:b + K_LANES]), src[b:b + K_LANES])
        return stats

    def isochrones(self, origins, limits, want_times=True):
        """Per-lane-limit isochrone sweeps: (reach bits [L, ceil(n/64)] u64, times [L, n] or None)."""
        org, lim = _u32(origins), _u32(limits)
        if org.size != lim.size:
            raise ValueError("origins and limits differ in length")
        words = (self.n_nodes + 63) // 64
        bits = np.empty((org.size, words), dtype=np.uint64)
        times = np.empty((org.size, self.n_nodes), dtype=np.uint32) if want_times else None
        _check(self._lib.bfg_isochrone(self._h, _native.ptr(org), _native.ptr(lim), org.size,
                                       _native.ptr(bits, _native.u64p),
                                       _native.ptr(times) if want_times else None))
        return bits, times

    @staticmethod
    def _seed_csr(seed_lists):
        L = len(seed_lists)
        off = np.zeros(L + 1, dtype=np.uint32)
        ranks, costs = [], []
        for i, seeds in enumerate(seed_lists):
            for r, c in seeds:
                ranks.append(r)
                costs.append(c)
            off[i + 1] = len(ranks)
        return off, _u32(ranks), _u32(costs)

    def isochrone_lists(self, seed_lists, thresholds, want_bits=False, capacity=None):
        """run_phast_bounded_fast_seeded (isochrone_handler.rs:338-491) for many
        lanes in one block-gated call: per lane an (m, 2) u32 array of
        (rank, dist <= threshold) in ascending rank, plus the reach bits
        [L, ceil(n/64)] when want_bits.  capacity: pairs to reserve (default
        an estimate; a second call runs when it is short)."""
        L = len(seed_lists)
        off, ranks, costs = self._seed_csr(seed_lists)
        thr = _u32(thresholds)
        if thr.size != L:
            raise ValueError("one threshold per lane")
        words = (self.n_nodes + 63) // 64
        bits = np.empty((L, words), dtype=np.uint64) if want_bits else None
        start = np.zeros(L, dtype=np.uint64)
        length = np.zeros(L, dtype=np.uint32)
        total = np.zeros(1, dtype=np.uint64)
        cap = int(capacity) if capacity is not None else min(L * self.n_nodes, max(1 << 16, L * 4096))
        P = _native.ptr
        for _ in range(2):
            pairs = np.empty(2 * max(cap, 1), dtype=np.uint32)
            rc = self._lib.bfg_isochrone_lists(self._h, P(off), P(ranks), P(costs), L, P(thr),
                                               P(bits, _native.u64p) if want_bits else None,
                                               P(start, _native.u64p), P(length), P(pairs), cap,
                                               P(total, _native.u64p))
            if rc == BFG_BUFFER_TOO_SMALL:
                cap = int(total[0])
                continue
            _check(rc)
            break
        pv = pairs.reshape(-1, 2)
        lists = [pv[int(s):int(s) + int(n)] for s, n in zip(start, length)]
        return (lists, bits) if want_bits else lists

    def isochrone_lists_device(self, origins, thresholds, d_bits, d_pairs, capacity):
        """Device-resident isochrone lists (bfg_isochrone_lists_device) of
        single-origin lanes: reach bits and (rank, dist) pairs stay in the HIP
        buffers d_bits / d_pairs (raw pointers); returns (start [L] u64,
        length [L] u32, total pairs, BFG result code: 0 or
        BFG_BUFFER_TOO_SMALL)."""
        org, thr = _u32(origins), _u32(thresholds)
        L = org.size
        if thr.size != L:
            raise ValueError("one threshold per lane")
        off = np.arange(L + 1, dtype=np.uint32)
        cost = np.zeros(L, dtype=np.uint32)
        start = np.zeros(L, dtype=np.uint64)
        length = np.zeros(L, dtype=np.uint32)
        total = np.zeros(1, dtype=np.uint64)
        P = _native.ptr
        rc = self._lib.bfg_isochrone_lists_device(self._h, P(off), P(org), P(cost), L, P(thr), d_bits,
                                                  P(start, _native.u64p), P(length), d_pairs, int(capacity),
                                                  P(total, _native.u64p))
        if rc not in (0, BFG_BUFFER_TOO_SMALL):
            _check(rc)
        return start, length, int(total[0]), rc

    def phast_seeded(self, seed_lists, thresholds):
        """Many run_phast_bounded_fast_seeded calls in one launch: dense rows [L, n]."""
        L = len(seed_lists)
        off = np.zeros(L + 1, dtype=np.uint32)
        ranks, costs = [], []
        for i, seeds in enumerate(seed_lists):
            for r, c in seeds:
                ranks.append(r)
                costs.append(c)
            off[i + 1] = len(ranks)
        ranks, costs, thr = _u32(ranks), _u32(costs), _u32(thresholds)
        if thr.size != L:
            raise ValueError("one threshold per lane")
        out = np.empty((L, self.n_nodes), dtype=np.uint32)
        _check(self._lib.bfg_phast_seeded(self._h, _native.ptr(off), _native.ptr(ranks), _native.ptr(costs), L,
                                          _native.ptr(thr), _native.ptr(out)))
        return out

    def attach_length(self, up_len, dn_len):
        """Length-along-time weights, one per CCH arc like the graph's up_w /
        dn_w (the reference's cch.lat weights, server/state.rs:2880-2891), for
        the two-channel calls."""
        up_len, dn_len = _u32(up_len), _u32(dn_len)
        if up_len.size != self.graph.m_up or dn_len.size != self.graph.m_down:
            raise ValueError("one length weight per arc")
        d = self.graph.desc()
        _check(self._lib.bfg_ctx_attach_length(self._h, C.byref(d), _native.ptr(up_len), _native.ptr(dn_len)))

    def phast_seeded_2ch(self, seed_lists, thresholds):
        """Many run_phast_bounded_fast_seeded_2ch calls (isochrone_handler.rs:
        500-625) in one launch; seeds are (rank, time, len).  Returns dense
        (times [L, n], lengths [L, n]), INF beyond each lane's limit."""
        L = len(seed_lists)
        off = np.zeros(L + 1, dtype=np.uint32)
        ranks, ts, ls = [], [], []
        for i, seeds in enumerate(seed_lists):
            for r, t, ln in seeds:
                ranks.append(r)
                ts.append(t)
                ls.append(ln)
            off[i + 1] = len(ranks)
        ranks, ts, ls, thr = _u32(ranks), _u32(ts), _u32(ls), _u32(thresholds)
        if thr.size != L:
            raise ValueError("one threshold per lane")
        times = np.empty((L, self.n_nodes), dtype=np.uint32)
        lens = np.empty((L, self.n_nodes), dtype=np.uint32)
        _check(self._lib.bfg_phast_seeded_2ch(self._h, _native.ptr(off), _native.ptr(ranks), _native.ptr(ts),
                                              _native.ptr(ls), L, _native.ptr(thr), _native.ptr(times),
                                              _native.ptr(lens)))
        return times, lens

    def run_phast_bounded_fast_seeded_2ch(self, seeds, threshold):
        """isochrone_handler.rs:500-509: Vec<(rank, time, len)> ascending rank, time <= threshold."""
        t, ln = self.phast_seeded_2ch([list(seeds)], [threshold])
        idx = np.nonzero(t[0] != INF)[0]
        return [(int(r), int(t[0][r]), int(ln[0][r])) for r in idx]

    def table_seeded_bounded_routed_2ch(self, src_seedsets, tgt_seedsets, threshold=INF):
        """bucket_ch.rs:6329-6379 on its PHAST path, table_phast_lopsided_2ch
        (:6381-6520).  Seeds are (rank, part_time, part_len, ...) tuples.
        Returns (S*T times, S*T lengths, S*T u8 conflicts), row-major; a
        conflict (a shared seed rank, which the reference recomputes with its
        two-channel bucket engine) is INF in both."""
        def flat(seedsets):
            off = np.zeros(len(seedsets) + 1, dtype=np.uint32)
            ranks, pt, pl = [], [], []
            for i, seeds in enumerate(seedsets):
                for sd in seeds:
                    ranks.append(sd[0])
                    pt.append(sd[1])
                    pl.append(sd[2])
                off[i + 1] = len(ranks)
            return off, _u32(ranks), _u32(pt), _u32(pl)
        so, sr, st, sl = flat(src_seedsets)
        to, tr, tt, tl = flat(tgt_seedsets)
        S, T = len(src_seedsets), len(tgt_seedsets)
        times = np.empty(S * T, dtype=np.uint32)
        lens = np.empty(S * T, dtype=np.uint32)
        conf = np.empty(S * T, dtype=np.uint8)
        P = _native.ptr
        _check(self._lib.bfg_table_seeded_2ch(self._h, P(so), P(sr), P(st), P(sl), S, P(to), P(tr), P(tt), P(tl),
                                              T, int(threshold) & INF, P(times), P(lens),
                                              conf.ctypes.data_as(C.POINTER(C.c_uint8))))
        return times, lens, conf

    def table_seeded_bounded_routed(self, src_seedsets, tgt_seedsets, threshold=INF):
        """bucket_ch.rs:6535-6574 on its PHAST path, table_phast_lopsided
        (:6717-6853): one seeded field per source, cell = min over the target's
        seeds (rank, part, ..) of F[rank] - part.  Seeds are (rank, part, ...)
        tuples (EngineSeed, :6039); only rank and part take part.  Returns
        (S*T u32 row-major, S*T u8 conflicts): a conflict is a pair whose
        source and target share a seed rank, which the reference recomputes
        with its bucket engine; those cells are INF here."""
        def flat(seedsets):
            off = np.zeros(len(seedsets) + 1, dtype=np.uint32)
            ranks, parts = [], []
            for i, seeds in enumerate(seedsets):
                for sd in seeds:
                    ranks.append(sd[0])
                    parts.append(sd[1])
                off[i + 1] = len(ranks)
            return off, _u32(ranks), _u32(parts)
        so, sr, sp = flat(src_seedsets)
        to, tr, tp = flat(tgt_seedsets)
        S, T = len(src_seedsets), len(tgt_seedsets)
        out = np.empty(S * T, dtype=np.uint32)
        conf = np.empty(S * T, dtype=np.uint8)
        _check(self._lib.bfg_table_seeded(self._h, _native.ptr(so), _native.ptr(sr), _native.ptr(sp), S,
                                          _native.ptr(to), _native.ptr(tr), _native.ptr(tp), T,
                                          int(threshold) & INF, _native.ptr(out),
                                          conf.ctypes.data_as(C.POINTER(C.c_uint8))))
        return out, conf

    def reverse_engine(self):
        """The engine over the reverse metric's CCH (the transposed CCH, built
        by the library), created on first use: a second device context whose
        rows are all-to-one distances
        (the reference passes the swapped flats, range/phast.rs:869-885)."""
        if self._rev is None:
            self._rev = BatchedPhastEngine(self.graph, device=self._device, transposed=True)
        return self._rev

    def pair_reverse(self):
        """Pairs this context with reverse_engine()'s (bfg_ctx_pair_reverse),
        the backward side of the bucket many-to-many; idempotent."""
        if not getattr(self, "_paired", False):
            rev = self.reverse_engine()
            _check(self._lib.bfg_ctx_pair_reverse(self._h, rev._h))
            self._paired = True
        return self._rev

    def table_bucket_parallel_bounded(self, sources, targets, threshold=INF, out=None):
        """bucket_ch.rs:3436-3520 on the GPU (bfg_table_bucket): the forward
        UP searches of the sources and the backward searches of the targets
        (upward phase on the transposed CCH) joined at their common
        elimination-tree ancestors -> (row-major S*T u32, stats); the cells
        equal compute_matrix_block_gated's."""
        self.pair_reverse()
        src, tgt = _u32(sources), _u32(targets)
        if out is None:
            out = np.empty(src.size * tgt.size, dtype=np.uint32)
        elif out.dtype != np.uint32 or out.size != src.size * tgt.size or not out.flags.c_contiguous:
            raise ValueError("out must be a contiguous u32 array of S*T cells")
        _check(self._lib.bfg_table_bucket(self._h, _native.ptr(src), src.size, _native.ptr(tgt), tgt.size,
                                          int(threshold) & INF, _native.ptr(out)))
        return out, self.last_stats(src.size)

    def table_bucket_parallel(self, sources, targets):
        """bucket_ch.rs:3436-3451: the unbounded bucket matrix."""
        return self.table_bucket_parallel_bounded(sources, targets, INF)

    def table_bucket_device(self, d_sources, n_src, d_targets, n_tgt, d_out, threshold=INF):
        """Device-resident bucket matrix (bfg_table_bucket_device): raw HIP pointers."""
        self.pair_reverse()
        _check(self._lib.bfg_table_bucket_device(self._h, d_sources, n_src, d_targets, n_tgt,
                                                 int(threshold) & INF, d_out))
        return self.last_stats(n_src)

    def query_block_gated_reverse(self, targets, threshold=INF):
        """PhastEngine::query_block_gated_reverse (range/phast.rs:879-1108) for
        many targets: dist[i][v] = d(v -> targets[i]), INF above threshold."""
        return self.reverse_engine().query_batch_bounded(targets, threshold)

    def run_phast_bounded_fast_seeded(self, seeds, threshold):
        """isochrone_handler.rs:338-344: Vec<(rank, dist)> ascending rank, dist <= threshold."""
        lst = self.isochrone_lists([list(seeds)], [threshold])[0]
        return [(int(r), int(d)) for r, d in lst]

    def run_phast_bounded_fast(self, origin_rank, threshold):
        """isochrone_handler.rs:315-329."""
        return self.run_phast_bounded_fast_seeded([(origin_rank, 0)], threshold)
