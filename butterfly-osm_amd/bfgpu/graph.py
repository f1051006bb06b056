"""The customized rank-aligned CCH on the host, and where it comes from.

`CchGraph` holds what the reference's `CchTopo` + `CchWeights` hold for the
path (route/src/formats/cch_topo.rs:250-286, cch_weights.rs:359-366):
up/down CSR by tail rank with u64 offsets, u32 targets and u32 weights
(0xFFFFFFFF = no arc).  `WeightArray` decoding (cch_weights.rs:115-187) turns
the reference's u16/u24/u32 storage into those u32 weights before upload.
"""
import ctypes as C
import dataclasses
import os

import numpy as np

from . import _native

INF = 0xFFFFFFFF
U24_SENTINEL = 0x00FFFFFF


# ---------------------------------------------------------------------------
# weight width codec (route/src/formats/cch_weights.rs:75-187)
# ---------------------------------------------------------------------------
def weight_width_choose(weights):
    """WeightWidth::choose: smallest of 'u16', 'u24', 'u32' that round-trips."""
    w = np.asarray(weights, dtype=np.uint64)
    fin = w[w != INF]
    if fin.size == 0 or fin.max() <= 0xFFFF - 1:
        return "u16"
    if fin.max() <= U24_SENTINEL - 1:
        return "u24"
    return "u32"


def weight_encode(weights, width):
    w = np.asarray(weights, dtype=np.uint32)
    if width == "u16":
        out = w.astype(np.uint16)
        out[w == INF] = 0xFFFF
        return out
    if width == "u24":
        v = np.where(w == INF, U24_SENTINEL, w).astype(np.uint32)
        b = np.empty((w.size, 3), dtype=np.uint8)
        b[:, 0] = v & 0xFF
        b[:, 1] = (v >> 8) & 0xFF
        b[:, 2] = (v >> 16) & 0xFF
        return b.reshape(-1)
    return w.copy()


def weight_decode(storage, width):
    """WeightArray::get for every entry: widen to u32, sentinel -> INF."""
    if width == "u16":
        s = np.asarray(storage, dtype=np.uint16)
        out = s.astype(np.uint32)
        out[s == 0xFFFF] = INF
        return out
    if width == "u24":
        b = np.asarray(storage, dtype=np.uint8).reshape(-1, 3).astype(np.uint32)
        v = b[:, 0] | (b[:, 1] << 8) | (b[:, 2] << 16)
        v[v == U24_SENTINEL] = INF
        return v
    if width == "u32":
        return np.asarray(storage, dtype=np.uint32).copy()
    raise ValueError(f"unknown weight width {width!r}")


@dataclasses.dataclass
class CchGraph:
    n: int
    up_off: np.ndarray
    up_tgt: np.ndarray
    up_w: np.ndarray
    dn_off: np.ndarray
    dn_tgt: np.ndarray
    dn_w: np.ndarray
    x: np.ndarray = None          # per-rank coordinates (metres), optional
    y: np.ndarray = None
    rank_to_filtered: np.ndarray = None
    ebg_off: np.ndarray = None    # edge-based graph in rank space (optional)
    ebg_head: np.ndarray = None
    ebg_w: np.ndarray = None
    meta: dict = dataclasses.field(default_factory=dict)

    def __post_init__(self):
        self.up_off = np.ascontiguousarray(self.up_off, dtype=np.uint64)
        self.dn_off = np.ascontiguousarray(self.dn_off, dtype=np.uint64)
        for f in ("up_tgt", "up_w", "dn_tgt", "dn_w"):
            setattr(self, f, np.ascontiguousarray(getattr(self, f), dtype=np.uint32))
        if self.up_off.size != self.n + 1 or self.dn_off.size != self.n + 1:
            raise ValueError("offset arrays must have n+1 entries")

    @property
    def m_up(self):
        return int(self.up_off[-1])

    @property
    def m_down(self):
        return int(self.dn_off[-1])

    def save(self, path):
        arrs = {k: v for k, v in dataclasses.asdict(self).items()
                if isinstance(v, np.ndarray)}
        np.savez(path, n=np.int64(self.n), **arrs)

    @staticmethod
    def load(path):
        z = np.load(path, allow_pickle=False)
        kw = {k: z[k] for k in z.files if k != "n"}
        return CchGraph(n=int(z["n"]), **kw)

    @staticmethod
    def from_arcs(n, up_arcs, down_arcs):
        """Builds the CSR from (tail, head, weight) lists, per-tail targets ascending
        (route/src/contraction.rs:893-912)."""
        def csr(arcs):
            arcs = sorted(arcs)
            off = np.zeros(n + 1, dtype=np.uint64)
            for t, _, _ in arcs:
                off[t + 1] += 1
            off = np.cumsum(off, dtype=np.uint64)
            tgt = np.array([h for _, h, _ in arcs], dtype=np.uint32)
            w = np.array([x for _, _, x in arcs], dtype=np.uint32)
            return off, tgt, w
        uo, ut, uw = csr(up_arcs)
        do, dt, dw = csr(down_arcs)
        return CchGraph(n=n, up_off=uo, up_tgt=ut, up_w=uw, dn_off=do, dn_tgt=dt, dn_w=dw)

    def transpose(self):
        """The CCH of the reverse metric: UP arcs are the DOWN arcs reversed
        (DownReverseAdjFlat, matrix/bucket_ch.rs:191-424, keyed by head) and DOWN
        arcs the UP arcs reversed, per-tail targets ascending.  A forward PHAST
        on it is the reference's reverse PHAST (range/phast.rs:879-1108:
        upward over DOWN arcs reversed, then a rank-descending pull over UP
        arcs): its label of v for origin t is d(v -> t)."""
        def flip(off, tgt, w):
            tails = np.repeat(np.arange(self.n, dtype=np.uint32), np.diff(off.astype(np.int64)))
            order = np.lexsort((tails, tgt))  # by new tail (old head), then new head
            noff = np.zeros(self.n + 1, dtype=np.uint64)
            noff[1:] = np.cumsum(np.bincount(tgt, minlength=self.n), dtype=np.uint64)
            return noff, tails[order], w[order]
        uo, ut, uw = flip(self.dn_off, self.dn_tgt, self.dn_w)
        do, dt, dw = flip(self.up_off, self.up_tgt, self.up_w)
        return CchGraph(n=self.n, up_off=uo, up_tgt=ut, up_w=uw, dn_off=do, dn_tgt=dt, dn_w=dw)

    def desc(self):
        """bfg_graph_desc pointing into this object's arrays (keep self alive)."""
        d = _native.GraphDesc()
        d.n_nodes = self.n
        d.up_offsets = _native.ptr(self.up_off, _native.u64p)
        d.up_targets = _native.ptr(self.up_tgt)
        d.up_weights = _native.ptr(self.up_w)
        d.down_offsets = _native.ptr(self.dn_off, _native.u64p)
        d.down_targets = _native.ptr(self.dn_tgt)
        d.down_weights = _native.ptr(self.dn_w)
        return d


@dataclasses.dataclass
class GridSpec:
    """Seeded synthetic road grid (the workload generator of the benchmarks)."""
    rows: int = 224
    cols: int = 224
    seed: int = 20261203
    spacing_m: float = 100.0
    jitter_frac: float = 0.25
    delete_frac: float = 0.10
    oneway_frac: float = 0.15
    chord_frac: float = 0.0
    ban_frac: float = 0.03
    leaf_threshold: int = 8192


def build_grid(spec: GridSpec, n_threads=0, verbose=False, keep_ebg=True) -> CchGraph:
    """Grid -> EBG -> ND order -> contraction -> customization (libbfcch)."""
    lib = _native.bfcch()
    p = _native.BfcParams(spec.rows, spec.cols, spec.seed, spec.spacing_m, spec.jitter_frac,
                          spec.delete_frac, spec.oneway_frac, spec.chord_frac, spec.ban_frac,
                          spec.leaf_threshold, n_threads, 1 if verbose else 0)
    h = lib.bfc_build_grid(C.byref(p))
    if not h:
        raise RuntimeError("cch build failed: " + lib.bfc_last_error().decode())
    try:
        c = h.contents
        n = c.n

        def arr(pt, count, dt):
            if count == 0:
                return np.zeros(0, dtype=dt)
            return np.ctypeslib.as_array(pt, shape=(count,)).astype(dt, copy=True)
        g = CchGraph(
            n=n,
            up_off=arr(c.up_off, n + 1, np.uint64), up_tgt=arr(c.up_tgt, c.m_up, np.uint32),
            up_w=arr(c.up_w, c.m_up, np.uint32),
            dn_off=arr(c.dn_off, n + 1, np.uint64), dn_tgt=arr(c.dn_tgt, c.m_down, np.uint32),
            dn_w=arr(c.dn_w, c.m_down, np.uint32),
            x=arr(c.x, n, np.float32), y=arr(c.y, n, np.float32),
            rank_to_filtered=arr(c.rank_to_filtered, n, np.uint32),
            meta=dict(nbg_nodes=c.nbg_nodes, nbg_edges=c.nbg_edges, n_components=c.n_components,
                      n_shortcuts=int(c.n_shortcuts), t_graph_s=c.t_graph_s, t_order_s=c.t_order_s,
                      t_contract_s=c.t_contract_s, spec=dataclasses.asdict(spec)))
        if keep_ebg:
            g.ebg_off = arr(c.ebg_off, n + 1, np.uint64)
            g.ebg_head = arr(c.ebg_head, c.m_ebg, np.uint32)
            g.ebg_w = arr(c.ebg_w, c.m_ebg, np.uint32)
        return g
    finally:
        lib.bfc_free(h)


def cached_grid(spec: GridSpec, cache_dir, **kw) -> CchGraph:
    """build_grid with an .npz cache keyed by the spec."""
    key = "grid_{rows}x{cols}_s{seed}_c{chord_frac}_b{ban_frac}_l{leaf_threshold}".format(**dataclasses.asdict(spec))
    path = os.path.join(cache_dir, key + ".npz")
    if os.path.exists(path):
        return CchGraph.load(path)
    g = build_grid(spec, **kw)
    os.makedirs(cache_dir, exist_ok=True)
    g.save(path)
    return g
