"""Host mirror of the reference's CCH customization on the GPU.

`triangle_relax_parallel` and `triangle_relax_lex_parallel` keep the names,
arguments and results of route/src/customization.rs:1083-1200 and
:1253-1440 (the topology, the arcs' weights before relaxation, the middles);
the relaxation runs in libbfgpu.so (`bfg_triangle_relax`,
`bfg_triangle_relax_lex`).  `original_arc_weights` gives the first weights of
a synthetic CCH whose builder did not keep contraction middles: the metric
weight of every original arc and INF for every shortcut, which relax to the
same fixpoint as the reference's bottom-up weights (a shortcut's bottom-up
weight is one of the triangle sums the relaxation takes the least of).
"""
import ctypes as C
import dataclasses

import numpy as np

from . import _native
from .engine import _check
from .graph import INF, CchGraph


@dataclasses.dataclass
class RelaxStats:
    """The call's counters: triangles evaluated, atomic mins that lowered an arc
    (schedule-dependent, like the reference's per-pass update counts),
    dependency levels, kernel launches, passes (always 1) and times."""
    triangles: int
    improvements: int
    levels: int
    launches: int
    passes: int
    kernel_ms: float
    total_ms: float


@dataclasses.dataclass
class RelaxResult:
    up_weights: np.ndarray
    down_weights: np.ndarray
    up_middles: np.ndarray
    down_middles: np.ndarray
    stats: RelaxStats


def _u32(a, size, what):
    a = np.ascontiguousarray(a, dtype=np.uint32)
    if a.size != size:
        raise ValueError(f"{what}: {a.size} values for {size} arcs")
    return a


def _bits(mask, size, what):
    mask = np.asarray(mask, dtype=bool)
    if mask.size != size:
        raise ValueError(f"{what}: {mask.size} flags for {size} arcs")
    words = np.zeros((size + 63) // 64 * 8, dtype=np.uint8)
    packed = np.packbits(mask, bitorder="little")
    words[:packed.size] = packed
    return words.view(np.uint64)


def _desc(topo: CchGraph, up_w, dn_w):
    d = _native.GraphDesc()
    d.n_nodes = topo.n
    d.up_offsets = _native.ptr(topo.up_off, _native.u64p)
    d.up_targets = _native.ptr(topo.up_tgt)
    d.up_weights = _native.ptr(up_w)
    d.down_offsets = _native.ptr(topo.dn_off, _native.u64p)
    d.down_targets = _native.ptr(topo.dn_tgt)
    d.down_weights = _native.ptr(dn_w)
    return d


def _stats(s):
    return RelaxStats(int(s.triangles), int(s.improvements), int(s.levels), int(s.launches), int(s.passes),
                      float(s.kernel_ms), float(s.total_ms))


def triangle_relax_parallel(topo: CchGraph, up_weights, down_weights, up_middles=None, down_middles=None,
                            device=0) -> RelaxResult:
    """triangle_relax_parallel (customization.rs:1083-1200) on the GPU: the
    relaxed (weights, middles) of every UP and DOWN arc of topo's CSR (its own
    weights are not read).  Middles default to u32::MAX (an original arc)."""
    mu, md = topo.m_up, topo.m_down
    uw = _u32(up_weights, mu, "up_weights")
    dw = _u32(down_weights, md, "down_weights")
    um = None if up_middles is None else _u32(up_middles, mu, "up_middles")
    dm = None if down_middles is None else _u32(down_middles, md, "down_middles")
    out = [np.empty(mu, np.uint32), np.empty(md, np.uint32), np.empty(mu, np.uint32), np.empty(md, np.uint32)]
    st = _native.RelaxStats()
    d = _desc(topo, uw, dw)
    P = _native.ptr
    _check(_native.bfgpu().bfg_triangle_relax(C.byref(d), int(device), None if um is None else P(um),
                                              None if dm is None else P(dm), *[P(a) for a in out], C.byref(st)))
    return RelaxResult(*out, _stats(st))


def triangle_relax_lex_parallel(topo: CchGraph, time_up, time_down, len_up, len_down, up_is_shortcut,
                                down_is_shortcut, up_middles=None, down_middles=None, device=0) -> RelaxResult:
    """triangle_relax_lex_parallel (customization.rs:1253-1440) on the GPU:
    relaxed times, and per shortcut the least apex whose halves sum to its
    converged (time, length) word (its given middle when none does)."""
    mu, md = topo.m_up, topo.m_down
    tu = _u32(time_up, mu, "time_up")
    td = _u32(time_down, md, "time_down")
    lu = _u32(len_up, mu, "len_up")
    ld = _u32(len_down, md, "len_down")
    su = _bits(up_is_shortcut, mu, "up_is_shortcut")
    sd = _bits(down_is_shortcut, md, "down_is_shortcut")
    um = None if up_middles is None else _u32(up_middles, mu, "up_middles")
    dm = None if down_middles is None else _u32(down_middles, md, "down_middles")
    out = [np.empty(mu, np.uint32), np.empty(md, np.uint32), np.empty(mu, np.uint32), np.empty(md, np.uint32)]
    st = _native.RelaxStats()
    d = _desc(topo, tu, td)
    P = _native.ptr
    _check(_native.bfgpu().bfg_triangle_relax_lex(C.byref(d), int(device), P(lu), P(ld), P(su, _native.u64p),
                                                  P(sd, _native.u64p), None if um is None else P(um),
                                                  None if dm is None else P(dm), *[P(a) for a in out],
                                                  C.byref(st)))
    return RelaxResult(*out, _stats(st))


def original_arc_weights(g: CchGraph):
    """(up_w0, down_w0, up_is_shortcut, down_is_shortcut) of g's CCH arcs from
    its rank-space edge-based graph (g.ebg_*): an arc that is an arc of the
    edge-based graph takes that arc's weight (the least, should there be
    several: the builder keeps the least), as compute_original_weight_rank_
    aligned gives it (customization.rs:1443-1470); every other arc is a
    shortcut and starts at INF."""
    if g.ebg_off is None:
        raise ValueError("the graph holds no edge-based arcs (built with keep_ebg=False)")
    n = np.uint64(g.n)
    tails = np.repeat(np.arange(g.n, dtype=np.uint64), np.diff(g.ebg_off.astype(np.int64)))
    heads = g.ebg_head.astype(np.uint64)
    res = []
    for off, tgt, sel in ((g.up_off, g.up_tgt, heads > tails), (g.dn_off, g.dn_tgt, heads < tails)):
        keys = np.repeat(np.arange(g.n, dtype=np.uint64), np.diff(off.astype(np.int64))) * n + tgt
        ek = tails[sel] * n + heads[sel]
        idx = np.searchsorted(keys, ek)
        if np.any(idx >= keys.size) or np.any(keys[np.minimum(idx, keys.size - 1)] != ek):
            raise ValueError("an edge-based arc is missing from the CCH topology")
        w0 = np.full(keys.size, INF, dtype=np.uint32)
        np.minimum.at(w0, idx, g.ebg_w[sel])
        orig = np.zeros(keys.size, dtype=bool)
        orig[idx] = True
        res.append((w0, ~orig))
    (uw, usc), (dw, dsc) = res
    return uw, dw, usc, dsc


def customize(g: CchGraph, device=0):
    """The graph with its weights customized on the GPU from its edge-based
    arcs (original weights, shortcuts INF, then triangle relaxation), and the
    relaxation's result."""
    uw, dw, _, _ = original_arc_weights(g)
    r = triangle_relax_parallel(g, uw, dw, device=device)
    out = dataclasses.replace(g, up_w=r.up_weights, dn_w=r.down_weights)
    return out, r
