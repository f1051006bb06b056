"""CPU tests of the host logic: C-ABI exports, layout/schedule validation, input
checks and the weight-width codec.  No device is touched."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from conftest import REPO, broom_graph, hand_graph

from bfgpu import CchGraph, _native, weight_decode, weight_encode, weight_width_choose
from bfgpu.graph import INF


def header_symbols():
    with open(os.path.join(REPO, "include", "bfgpu.h")) as f:
        txt = f.read()
    return sorted(set(re.findall(r"^\s*(?:int|void|const char \*)\s*\*?\s*(bfg_\w+)\s*\(", txt, re.M)))


def test_library_exports_every_declared_symbol():
    syms = header_symbols()
    assert len(syms) >= 12
    lib = _native.bfgpu()  # loads without a device
    for s in syms:
        assert hasattr(lib, s), s
        assert s in _native.BFGPU_SYMBOLS, f"{s} has no ctypes binding"


def validate(g):
    info = _native.CtxInfo()
    d = g.desc()
    rc = _native.bfgpu().bfg_validate_graph(C.byref(d), C.byref(info))
    return rc, info


@pytest.mark.parametrize("name", ["five_node", "asymmetric_three_node", "inf_arc_four_node"])
def test_validate_hand_graphs(name):
    g, _ = hand_graph(name)
    rc, info = validate(g)
    assert rc == 0, _native.bfgpu().bfg_last_error()
    assert info.n_nodes == g.n
    if name == "inf_arc_four_node":
        assert (info.m_up, info.m_down) == (3, 3)  # INF arcs dropped


def test_validate_broom_and_grids(grid_small, grid_config1):
    g, _, _ = broom_graph(1000)
    rc, info = validate(g)
    assert rc == 0 and info.etree_depth == 1
    for gg in (grid_small, grid_config1):
        rc, info = validate(gg)
        assert rc == 0, _native.bfgpu().bfg_last_error()
        assert info.m_down == gg.m_down and info.m_up == gg.m_up
        assert info.n_bottom_levels >= 1
        assert info.n_segments >= info.n_segment_levels
        assert 0 < info.n_top < gg.n


def test_rejects_non_rank_aligned_cch():
    g = CchGraph.from_arcs(3, [(2, 1, 5)], [])  # UP arc going down
    rc, _ = validate(g)
    assert rc == 1
    assert b"UP arc" in _native.bfgpu().bfg_last_error()
    g = CchGraph.from_arcs(3, [], [(0, 2, 5)])  # DOWN arc going up
    assert validate(g)[0] == 1


def test_weight_width_codec():
    # WeightWidth::choose boundaries (cch_weights.rs:106-132)
    assert weight_width_choose([0, 65534, INF]) == "u16"
    assert weight_width_choose([65535]) == "u24"
    assert weight_width_choose([0xFFFFFE]) == "u24"
    assert weight_width_choose([0xFFFFFF]) == "u32"
    rng = np.random.default_rng(1)
    for hi in (65534, 0xFFFFFE, 0xFFFFFFFE):
        w = rng.integers(0, hi + 1, 1000, dtype=np.uint64).astype(np.uint32)
        w[::7] = INF
        width = weight_width_choose(w)
        enc = weight_encode(w, width)
        assert enc.dtype.itemsize * (3 if width == "u24" else 1) == {"u16": 2, "u24": 3, "u32": 4}[width] * (
            3 if width == "u24" else 1) or width == "u24"
        assert np.array_equal(weight_decode(enc, width), w)
    # sentinel of each width widens to u32::MAX (WeightArray::get)
    assert weight_decode(np.array([0xFFFF, 7], dtype=np.uint16), "u16").tolist() == [INF, 7]
    assert weight_decode(np.array([0xFF, 0xFF, 0xFF, 1, 0, 0], dtype=np.uint8), "u24").tolist() == [INF, 1]


def test_transpose_is_an_involution_and_keeps_rank_alignment(grid_small):
    """CchGraph.transpose (the reverse metric's CCH): UP targets stay above their
    tail, DOWN targets below, and transposing twice gives the CCH back."""
    g = grid_small
    r = g.transpose()
    tails_up = np.repeat(np.arange(g.n), np.diff(r.up_off.astype(np.int64)))
    tails_dn = np.repeat(np.arange(g.n), np.diff(r.dn_off.astype(np.int64)))
    assert (r.up_tgt > tails_up).all() and (r.dn_tgt < tails_dn).all()
    assert r.m_up == g.m_down and r.m_down == g.m_up
    rr = r.transpose()
    for f in ("up_off", "up_tgt", "up_w", "dn_off", "dn_tgt", "dn_w"):
        a, b = getattr(g, f), getattr(rr, f)
        if f.endswith("tgt") or f.endswith("_w"):  # per-tail order may differ: compare sorted (tail, head, w)
            continue
        assert np.array_equal(a, b), f
    for off, tgt, w, off2, tgt2, w2 in ((g.up_off, g.up_tgt, g.up_w, rr.up_off, rr.up_tgt, rr.up_w),
                                        (g.dn_off, g.dn_tgt, g.dn_w, rr.dn_off, rr.dn_tgt, rr.dn_w)):
        t = np.repeat(np.arange(g.n), np.diff(off.astype(np.int64)))
        k1 = np.lexsort((w, tgt, t))
        k2 = np.lexsort((w2, tgt2, t))
        assert np.array_equal(tgt[k1], tgt2[k2]) and np.array_equal(w[k1], w2[k2])



def _enc16(x, cap=0x7FFF):
    """The 16-bit label and weight encoding of the bucket's pushes and join:
    0xFFFF for INF, else min(x, cap)."""
    x = np.asarray(x, dtype=np.uint64)
    return np.where(x == 0xFFFFFFFF, 0xFFFF, np.minimum(x, cap)).astype(np.uint64)


def test_16bit_cap_encoding_min_plus():
    """The property the 16-bit pushes and join rely on, on random min-plus
    products (labels with INF, values on both sides of the cap): with
    saturating 16-bit sums of encoded operands, a result of 0xFFFF is INF
    exactly when the u32 result is, a result below the cap equals the u32
    result, and a result in [cap, 0xFFFE] stands for a finite u32 result of
    at least the cap (the case the kernels recompute in u32).  The u32
    reference is the oracle's arithmetic: saturating adds, INF = 2^32 - 1."""
    rng = np.random.default_rng(16)
    INF32 = np.uint64(0xFFFFFFFF)
    for hi in (0x3000, 0x7FFF, 0x9000, 0x30000):
        a = rng.integers(0, hi, (400, 24)).astype(np.uint64)
        b = rng.integers(0, hi, (400, 24)).astype(np.uint64)
        a[rng.random(a.shape) < 0.2] = INF32
        b[rng.random(b.shape) < 0.2] = INF32
        a[:5] = INF32  # rows with every sum INF
        s32 = np.where((a == INF32) | (b == INF32), INF32, np.minimum(a + b, INF32))
        r32 = s32.min(axis=1)
        r16 = np.minimum(_enc16(a) + _enc16(b), 0xFFFF).min(axis=1)
        assert np.array_equal(r16 == 0xFFFF, r32 == INF32)
        low = r16 < 0x7FFF
        assert np.array_equal(r16[low], r32[low])
        mid = (r16 >= 0x7FFF) & (r16 < 0xFFFF)
        assert np.all((r32[mid] >= 0x7FFF) & (r32[mid] != INF32))
        # a u32 result below the cap is never reported in the recompute range
        assert np.all(r16[r32 < 0x7FFF] == r32[r32 < 0x7FFF])


def test_16bit_per_cell_skip_rule():
    """The 16-bit join's per-cell bound: a segment is skipped when, for every
    cell of a block, the saturating 16-bit sum of its two lanes' encoded
    least labels in the segment is no less than the cell.  Whenever that
    holds, relaxing the segment (the 16-bit min-plus over its positions)
    changes no cell below the cap and leaves an INF cell INF; cells at or
    past the cap are recomputed in u32 by the kernels whatever happens."""
    rng = np.random.default_rng(61)
    INF32 = 0xFFFFFFFF
    skipped = 0
    for trial in range(3000):
        hi = (0x2000, 0x8000, 0x20000)[trial % 3]
        A = rng.integers(0, hi, (8, 64)).astype(np.uint64)
        B = rng.integers(0, hi, (8, 64)).astype(np.uint64)
        A[rng.random(A.shape) < 0.1] = INF32
        B[rng.random(B.shape) < 0.1] = INF32
        if trial % 7 == 0:
            A[rng.integers(0, 8)] = INF32  # a lane that does not reach the segment
        cand = np.minimum(_enc16(A.min(axis=1))[:, None] + _enc16(B.min(axis=1))[None, :], 0xFFFF)
        # cells so far at or a little below the bound (and above it now and then)
        cells = cand.astype(np.int64) - rng.integers(-3, 60, cand.shape)
        cells = np.clip(cells, 0, 0xFFFF).astype(np.uint64)
        if not np.all(cand >= cells):
            continue
        skipped += 1
        sums = np.minimum(_enc16(A)[:, None, :] + _enc16(B)[None, :, :], 0xFFFF).min(axis=2)
        new = np.minimum(cells, sums)
        kept = (cells < 0x7FFF) | (cells == 0xFFFF)
        assert np.array_equal(new[kept], cells[kept])
    assert skipped > 100
