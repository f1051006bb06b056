"""ctypes wrapper of liborc.so — the CPU restatement of the reference's PHAST.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, always as the checker or the timed CPU
reference, never by the product path.  Each function names the reference code
it restates (see phast_oracle.c for the file:line table).
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
INF = 0xFFFFFFFF
K_LANES = 8

u32p = C.POINTER(C.c_uint32)
u64p = C.POINTER(C.c_uint64)


class _Cch(C.Structure):
    _fields_ = [("n", C.c_uint32), ("up_off", u64p), ("up_tgt", u32p), ("up_w", u32p),
                ("dn_off", u64p), ("dn_tgt", u32p), ("dn_w", u32p)]


class _Flat(C.Structure):
    _fields_ = [("n", C.c_uint32), ("off", u64p), ("tgt", u32p), ("w", u32p)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "liborc.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is missing: run oracle/build.py (or __graft_entry__.build())")
        _lib = C.CDLL(path)
        cp = C.POINTER(_Cch)
        fp = C.POINTER(_Flat)
        sig = {
            "orc_query_batch": [cp, u32p, C.c_uint32, u32p],
            "orc_query_batch_soa": [cp, u32p, C.c_uint32, u32p],
            "orc_query_batch_bounded": [cp, u32p, C.c_uint32, C.c_uint32, u32p],
            "orc_query_batch_block_gated": [cp, u32p, C.c_uint32, C.c_uint32, u32p],
            "orc_compute_matrix_flat": [cp, u32p, C.c_uint32, u32p, C.c_uint32, u32p],
            "orc_compute_matrix_flat_soa": [cp, u32p, C.c_uint32, u32p, C.c_uint32, u32p],
            "orc_compute_matrix_block_gated": [cp, u32p, C.c_uint32, u32p, C.c_uint32, C.c_uint32, u32p],
            "orc_compute_matrix_flat_mt": [cp, u32p, C.c_uint32, u32p, C.c_uint32, u32p, C.c_int],
            "orc_phast_query": [cp, C.c_uint32, u32p],
            "orc_phast_query_reverse": [cp, C.c_uint32, u32p],
            "orc_build_up_flat": [cp, fp],
            "orc_build_down_flat": [cp, fp],
            "orc_build_down_reverse_flat": [cp, fp],
            "orc_dijkstra": [C.c_uint32, u64p, u32p, u32p, C.c_uint32, u32p],
            "orc_table_bucket": [fp, fp, u32p, C.c_uint32, u32p, C.c_uint32, C.c_uint32, u32p],
        }
        for k, v in sig.items():
            f = getattr(_lib, k)
            f.argtypes = v
            f.restype = C.c_int
        _lib.orc_triangle_relax.argtypes = [C.c_uint32, u64p, u32p, u64p, u32p, u64p, u64p, C.c_int, u64p]
        _lib.orc_triangle_relax.restype = C.c_int
        _lib.orc_elect_middles.argtypes = [C.c_uint32, u64p, u32p, u64p, u32p, u64p, u64p, u64p, u64p, u32p, u32p,
                                           u32p, u32p]
        _lib.orc_elect_middles.restype = C.c_int
        _lib.orc_flat_free.argtypes = [fp]
        _lib.orc_flat_free.restype = None
        _lib.orc_seeded_state_new.argtypes = [C.c_uint32]
        _lib.orc_seeded_state_new.restype = C.c_void_p
        _lib.orc_seeded_state_free.argtypes = [C.c_void_p]
        _lib.orc_seeded_state_free.restype = None
        _lib.orc_phast_seeded.argtypes = [C.c_void_p, fp, fp, u32p, u32p, C.c_uint32, C.c_uint32, u32p, u32p]
        _lib.orc_phast_seeded.restype = C.c_uint64
        _lib.orc_phast_seeded_2ch.argtypes = [C.c_void_p, fp, fp, fp, fp, u32p, u32p, u32p, C.c_uint32,
                                              C.c_uint32, u32p, u32p, u32p]
        _lib.orc_phast_seeded_2ch.restype = C.c_uint64
    return _lib


def _p(a, t=u32p):
    return a.ctypes.data_as(t)


def _u32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.uint32))


class Oracle:
    """Holds one CCH (anything with n, up_off, up_tgt, up_w, dn_off, dn_tgt, dn_w)."""

    def __init__(self, g):
        self.g = g
        self.n = int(g.n)
        self._arrs = [np.ascontiguousarray(g.up_off, dtype=np.uint64), _u32(g.up_tgt), _u32(g.up_w),
                      np.ascontiguousarray(g.dn_off, dtype=np.uint64), _u32(g.dn_tgt), _u32(g.dn_w)]
        a = self._arrs
        self._c = _Cch(self.n, _p(a[0], u64p), _p(a[1]), _p(a[2]), _p(a[3], u64p), _p(a[4]), _p(a[5]))
        self.L = lib()

    # batched_phast.rs:131-260 (k <= 8)
    def query_batch(self, sources):
        s = _u32(sources)
        assert s.size <= K_LANES, "Too many sources for batch"
        out = np.empty((s.size, self.n), dtype=np.uint32)
        self.L.orc_query_batch(C.byref(self._c), _p(s), s.size, _p(out))
        return out

    # batched_phast.rs:800-917, returned as AoS
    def query_batch_soa(self, sources):
        s = _u32(sources)
        assert s.size <= K_LANES
        ds = np.empty(self.n * K_LANES, dtype=np.uint32)
        self.L.orc_query_batch_soa(C.byref(self._c), _p(s), s.size, _p(ds))
        return ds.reshape(self.n, K_LANES)[:, :s.size].T.copy()

    def query_batch_bounded(self, sources, threshold):
        s = _u32(sources)
        assert s.size <= K_LANES
        out = np.empty((s.size, self.n), dtype=np.uint32)
        self.L.orc_query_batch_bounded(C.byref(self._c), _p(s), s.size, threshold, _p(out))
        return out

    def query_batch_block_gated(self, sources, threshold):
        s = _u32(sources)
        assert s.size <= K_LANES
        out = np.empty((s.size, self.n), dtype=np.uint32)
        self.L.orc_query_batch_block_gated(C.byref(self._c), _p(s), s.size, threshold, _p(out))
        return out

    def compute_matrix_flat(self, sources, targets):
        s, t = _u32(sources), _u32(targets)
        out = np.empty(s.size * t.size, dtype=np.uint32)
        rc = self.L.orc_compute_matrix_flat(C.byref(self._c), _p(s), s.size, _p(t), t.size, _p(out))
        assert rc == 0
        return out

    def compute_matrix_flat_soa(self, sources, targets):
        s, t = _u32(sources), _u32(targets)
        out = np.empty(s.size * t.size, dtype=np.uint32)
        assert self.L.orc_compute_matrix_flat_soa(C.byref(self._c), _p(s), s.size, _p(t), t.size, _p(out)) == 0
        return out

    def compute_matrix_block_gated(self, sources, targets, threshold):
        s, t = _u32(sources), _u32(targets)
        out = np.empty(s.size * t.size, dtype=np.uint32)
        assert self.L.orc_compute_matrix_block_gated(C.byref(self._c), _p(s), s.size, _p(t), t.size, threshold,
                                                     _p(out)) == 0
        return out

    def compute_matrix_flat_mt(self, sources, targets, n_threads=0):
        s, t = _u32(sources), _u32(targets)
        out = np.empty(s.size * t.size, dtype=np.uint32)
        assert self.L.orc_compute_matrix_flat_mt(C.byref(self._c), _p(s), s.size, _p(t), t.size, _p(out),
                                                 n_threads) == 0
        return out

    def phast_query(self, origin):
        out = np.empty(self.n, dtype=np.uint32)
        self.L.orc_phast_query(C.byref(self._c), int(origin), _p(out))
        return out

    def phast_query_reverse(self, target):
        out = np.empty(self.n, dtype=np.uint32)
        self.L.orc_phast_query_reverse(C.byref(self._c), int(target), _p(out))
        return out

    def _flat(self, which):
        f = _Flat()
        getattr(self.L, which)(C.byref(self._c), C.byref(f))
        m = int(f.off[self.n])
        res = (np.ctypeslib.as_array(f.off, shape=(self.n + 1,)).copy(),
               np.ctypeslib.as_array(f.tgt, shape=(max(m, 1),))[:m].copy(),
               np.ctypeslib.as_array(f.w, shape=(max(m, 1),))[:m].copy())
        self.L.orc_flat_free(C.byref(f))
        return res

    def up_flat(self):
        return self._flat("orc_build_up_flat")

    def down_flat(self):
        return self._flat("orc_build_down_flat")

    def down_reverse_flat(self):
        return self._flat("orc_build_down_reverse_flat")

    def _flat_struct(self, which):
        """An INF-filtered flat as the C struct (kept alive on this object)."""
        if not hasattr(self, "_flats"):
            self._flats = {}
        if which not in self._flats:
            off, tgt, w = self._flat(which)
            arrs = (np.ascontiguousarray(off, dtype=np.uint64), _u32(tgt), _u32(w))
            f = _Flat(self.n, arrs[0].ctypes.data_as(u64p), _p(arrs[1]), _p(arrs[2]))
            self._flats[which] = (f, arrs)
        return self._flats[which][0]

    def table_bucket(self, sources, targets, threshold=INF):
        """table_bucket_parallel_bounded (bucket_ch.rs:3436-3520): the bucket
        many-to-many, row-major S*T; cells above threshold are unspecified
        (the reference's callers mask them)."""
        s, t = _u32(sources), _u32(targets)
        out = np.empty(s.size * t.size, dtype=np.uint32)
        up = self._flat_struct("orc_build_up_flat")
        dr = self._flat_struct("orc_build_down_reverse_flat")
        self.L.orc_table_bucket(C.byref(up), C.byref(dr), _p(s), s.size, _p(t), t.size, int(threshold) & INF,
                                _p(out))
        return out

    def seeded(self, seeds, threshold):
        """run_phast_bounded_fast_seeded -> list of (rank, dist)."""
        if not hasattr(self, "_st"):
            self._st = self.L.orc_seeded_state_new(self.n)
            uo, ut, uw = self.up_flat()
            do, dt, dw = self.down_flat()
            self._flats = [uo, ut, uw, do, dt, dw]
            self._fu = _Flat(self.n, _p(uo, u64p), _p(ut), _p(uw))
            self._fd = _Flat(self.n, _p(do, u64p), _p(dt), _p(dw))
        r = _u32([s[0] for s in seeds])
        c = _u32([s[1] for s in seeds])
        orank = np.empty(self.n, dtype=np.uint32)
        odist = np.empty(self.n, dtype=np.uint32)
        cnt = self.L.orc_phast_seeded(self._st, C.byref(self._fu), C.byref(self._fd), _p(r), _p(c), r.size,
                                      threshold, _p(orank), _p(odist))
        return list(zip(orank[:cnt].tolist(), odist[:cnt].tolist()))

    # matrix/bucket_ch.rs:6717-6853 (table_phast_lopsided), the bucket
    # recomputation of same-edge conflicts (:6828-6845) left out: returns
    # (S*T cells, S*T conflict flags), a conflicted cell INF
    def table_phast_lopsided(self, src_seedsets, tgt_seedsets, threshold):
        S, T = len(src_seedsets), len(tgt_seedsets)
        cells = np.full(S * T, INF, dtype=np.uint32)
        conf = np.zeros(S * T, dtype=np.uint8)
        # :6738-6752 the field bound widens the threshold by both sides' worst parts
        max_tgt_part = max((s[1] for v in tgt_seedsets for s in v), default=0)

        def field_bound(seeds):
            if threshold == INF:
                return INF
            return min(INF, min(INF, threshold + max((s[1] for s in seeds), default=0)) + max_tgt_part)

        tgt_ranks = {s[0] for v in tgt_seedsets for s in v}
        for i, seeds in enumerate(src_seedsets):  # :6753
            if not seeds:  # :6756-6758
                continue
            src_ranks = {s[0] for s in seeds}
            settled = self.seeded([(s[0], s[1]) for s in seeds], field_bound(seeds))  # :6762-6769
            field = {}
            for rank, dist in settled:  # :6771-6783
                if rank in tgt_ranks:
                    field[rank] = min(dist, field.get(rank, INF))
            for j, tseeds in enumerate(tgt_seedsets):  # :6785-6811
                if not tseeds:
                    continue
                if any(s[0] in src_ranks for s in tseeds):  # :6789-6793
                    conf[i * T + j] = 1
                    continue
                best = INF
                for s in tseeds:
                    f = field.get(s[0])
                    if f is not None and f != INF:
                        best = min(best, max(0, f - s[1]))  # saturating_sub
                if threshold != INF and best != INF and best > threshold:
                    best = INF
                cells[i * T + j] = best
        return cells, conf

    def attach_length(self, up_len, dn_len):
        """Length-along-time weights, one per CCH arc like up_w / dn_w.  The
        reference builds their flats from the length weights
        (UpAdjFlat::build / DownAdjFlat::build, server/state.rs:178-183,
        :2890-2891) and indexes them with the time flats' arc index, so both
        must drop the same arcs."""
        self.seeded([], INF)  # builds the time flats and the state
        lg = _Cch(self.n, self._c.up_off, self._c.up_tgt, _p(_u32(up_len)), self._c.dn_off, self._c.dn_tgt,
                  _p(_u32(dn_len)))
        self._len_src = [_u32(up_len), _u32(dn_len)]
        lg.up_w, lg.dn_w = _p(self._len_src[0]), _p(self._len_src[1])
        saved = self._c
        self._c = lg
        try:
            luo, lut, luw = self.up_flat()
            ldo, ldt, ldw = self.down_flat()
        finally:
            self._c = saved
        uo, ut, _, do, dt, _ = self._flats
        if not (np.array_equal(luo, uo) and np.array_equal(lut, ut) and np.array_equal(ldo, do)
                and np.array_equal(ldt, dt)):
            raise ValueError("length flats do not share the time flats' topology")
        self._lflats = [luw, ldw]
        self._flu = _Flat(self.n, _p(uo, u64p), _p(ut), _p(luw))
        self._fld = _Flat(self.n, _p(do, u64p), _p(dt), _p(ldw))

    def seeded_2ch(self, seeds, threshold):
        """run_phast_bounded_fast_seeded_2ch (isochrone_handler.rs:500-625):
        seeds (rank, time, len) -> list of (rank, time, len), time <= threshold."""
        r = _u32([s[0] for s in seeds])
        t = _u32([s[1] for s in seeds])
        ln = _u32([s[2] for s in seeds])
        orank = np.empty(self.n, dtype=np.uint32)
        otime = np.empty(self.n, dtype=np.uint32)
        olen = np.empty(self.n, dtype=np.uint32)
        cnt = self.L.orc_phast_seeded_2ch(self._st, C.byref(self._fu), C.byref(self._fd), C.byref(self._flu),
                                          C.byref(self._fld), _p(r), _p(t), _p(ln), r.size, threshold,
                                          _p(orank), _p(otime), _p(olen))
        return list(zip(orank[:cnt].tolist(), otime[:cnt].tolist(), olen[:cnt].tolist()))

    def seeded_2ch_rows(self, seeds, threshold):
        """seeded_2ch as dense (time, len) rows over all ranks, INF where unsettled."""
        tr = np.full(self.n, INF, dtype=np.uint32)
        lr = np.full(self.n, INF, dtype=np.uint32)
        for rank, t, ln in self.seeded_2ch(seeds, threshold):
            tr[rank] = t
            lr[rank] = ln
        return tr, lr

    # matrix/bucket_ch.rs:6381-6520 (table_phast_lopsided_2ch), the bucket
    # recomputation of same-edge conflicts (:6496-6516) left out: returns
    # (S*T times, S*T lengths, S*T conflict flags), a conflicted pair INF in both
    def table_phast_lopsided_2ch(self, src_seedsets, tgt_seedsets, threshold):
        S, T = len(src_seedsets), len(tgt_seedsets)
        tm = np.full(S * T, INF, dtype=np.uint32)
        lm = np.full(S * T, INF, dtype=np.uint32)
        conf = np.zeros(S * T, dtype=np.uint8)
        max_tgt_part = max((s[1] for v in tgt_seedsets for s in v), default=0)

        def field_bound(seeds):  # :6407-6415
            if threshold == INF:
                return INF
            return min(INF, min(INF, threshold + max((s[1] for s in seeds), default=0)) + max_tgt_part)

        tgt_ranks = {s[0] for v in tgt_seedsets for s in v}
        for i, seeds in enumerate(src_seedsets):
            if not seeds:
                continue
            src_ranks = {s[0] for s in seeds}
            settled = self.seeded_2ch([(s[0], s[1], s[2]) for s in seeds], field_bound(seeds))  # :6431-6439
            field = {}
            for rank, t, ln in settled:  # :6440-6453
                if rank in tgt_ranks and (rank not in field or t < field[rank][0]):
                    field[rank] = (t, ln)
            for j, tseeds in enumerate(tgt_seedsets):  # :6454-6481
                if not tseeds:
                    continue
                if any(s[0] in src_ranks for s in tseeds):
                    conf[i * T + j] = 1
                    continue
                bt, bl = INF, INF
                for s in tseeds:
                    f = field.get(s[0])
                    if f is not None and f[0] != INF:
                        at = max(0, f[0] - s[1])
                        if at < bt:  # strict: the first target seed of the best time wins
                            bt, bl = at, max(0, f[1] - s[2])
                if threshold != INF and bt != INF and bt > threshold:
                    bt, bl = INF, INF
                tm[i * T + j] = bt
                lm[i * T + j] = bl
        return tm, lm, conf

    def __del__(self):
        st = getattr(self, "_st", None)
        if st:
            self.L.orc_seeded_state_free(st)
            self._st = None


def _pack(hi, lo):
    return (np.asarray(hi, dtype=np.uint64) << np.uint64(32)) | np.asarray(lo, dtype=np.uint64)


def _sc_words(mask):
    mask = np.asarray(mask, dtype=bool)
    words = np.zeros((mask.size + 63) // 64 * 8, dtype=np.uint8)
    b = np.packbits(mask, bitorder="little")
    words[:b.size] = b
    return words.view(np.uint64)


def triangle_relax(g, up_w, dn_w, up_mid=None, dn_mid=None):
    """triangle_relax_parallel (customization.rs:1083-1200) over g's topology:
    (up weights, down weights, up middles, down middles, passes)."""
    up_pk = _pack(up_w, INF if up_mid is None else up_mid)
    dn_pk = _pack(dn_w, INF if dn_mid is None else dn_mid)
    passes = lib().orc_triangle_relax(g.n, _p(g.up_off, u64p), _p(g.up_tgt), _p(g.dn_off, u64p), _p(g.dn_tgt),
                                      _p(up_pk, u64p), _p(dn_pk, u64p), 0, None)
    if passes < 0:
        raise RuntimeError("CCH customization did not converge after 100 passes")
    lo = np.uint64(0xFFFFFFFF)
    return ((up_pk >> np.uint64(32)).astype(np.uint32), (dn_pk >> np.uint64(32)).astype(np.uint32),
            (up_pk & lo).astype(np.uint32), (dn_pk & lo).astype(np.uint32), passes)


def triangle_relax_lex(g, t_up, t_dn, len_up, len_dn, up_sc, dn_sc, up_mid=None, dn_mid=None):
    """triangle_relax_lex_parallel (customization.rs:1253-1440) over g's
    topology: (up times, down times, up middles, down middles, passes, up
    lengths, down lengths), the lengths being the converged low words."""
    up_pk = _pack(t_up, len_up)
    dn_pk = _pack(t_dn, len_dn)
    L = lib()
    passes = L.orc_triangle_relax(g.n, _p(g.up_off, u64p), _p(g.up_tgt), _p(g.dn_off, u64p), _p(g.dn_tgt),
                                  _p(up_pk, u64p), _p(dn_pk, u64p), 1, None)
    um = np.empty(g.m_up, dtype=np.uint32)
    dm = np.empty(g.m_down, dtype=np.uint32)
    umi = None if up_mid is None else _u32(up_mid)
    dmi = None if dn_mid is None else _u32(dn_mid)
    L.orc_elect_middles(g.n, _p(g.up_off, u64p), _p(g.up_tgt), _p(g.dn_off, u64p), _p(g.dn_tgt), _p(up_pk, u64p),
                        _p(dn_pk, u64p), _p(_sc_words(up_sc), u64p), _p(_sc_words(dn_sc), u64p),
                        None if umi is None else _p(umi), None if dmi is None else _p(dmi), _p(um), _p(dm))
    lo = np.uint64(0xFFFFFFFF)
    return ((up_pk >> np.uint64(32)).astype(np.uint32), (dn_pk >> np.uint64(32)).astype(np.uint32), um, dm, passes,
            (up_pk & lo).astype(np.uint32), (dn_pk & lo).astype(np.uint32))


def dijkstra(n, off, head, w, src):
    out = np.empty(int(n), dtype=np.uint32)
    off = np.ascontiguousarray(off, dtype=np.uint64)
    head, w = _u32(head), _u32(w)
    lib().orc_dijkstra(int(n), _p(off, u64p), _p(head), _p(w), int(src), _p(out))
    return out


def union_graph(g):
    """CSR of the CCH union graph (UP and DOWN arcs by tail), for brute force."""
    n = int(g.n)
    up_cnt = np.diff(np.asarray(g.up_off, dtype=np.int64))
    dn_cnt = np.diff(np.asarray(g.dn_off, dtype=np.int64))
    off = np.zeros(n + 1, dtype=np.uint64)
    off[1:] = np.cumsum(up_cnt + dn_cnt)
    head = np.empty(int(off[-1]), dtype=np.uint32)
    w = np.empty(int(off[-1]), dtype=np.uint32)
    uo = np.asarray(g.up_off, dtype=np.int64)
    do = np.asarray(g.dn_off, dtype=np.int64)
    for r in range(n):
        a = int(off[r])
        k = int(up_cnt[r])
        head[a:a + k] = g.up_tgt[uo[r]:uo[r + 1]]
        w[a:a + k] = g.up_w[uo[r]:uo[r + 1]]
        k2 = int(dn_cnt[r])
        head[a + k:a + k + k2] = g.dn_tgt[do[r]:do[r + 1]]
        w[a + k:a + k + k2] = g.dn_w[do[r]:do[r + 1]]
    return off, head, w
