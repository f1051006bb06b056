import json
import os
import sys

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(REPO, "butterfly-osm_amd")
for p in (REPO, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(REPO, "tests", "golden")
CACHE = os.environ.get("BFG_TEST_CACHE", "/tmp/bfg_test_cache")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) to run")


def _ensure_built():
    """Build only what is missing (a fresh checkout builds everything here)."""
    if not os.path.exists(os.path.join(REPO, "oracle", "liborc.so")):
        from oracle import build as ob
        ob.build()
    need = [os.path.join(PKG, "lib", n) for n in ("libbfgpu.so", "libbfcch.so")]
    if not all(os.path.exists(p) for p in need):
        sys.path.insert(0, PKG)
        import build as pb
        pb.build_all()


_ensure_built()


def hand_vectors():
    with open(os.path.join(GOLDEN, "hand_vectors.json")) as f:
        return json.load(f)


def hand_graph(name):
    from bfgpu import CchGraph
    hv = hand_vectors()[name]
    g = CchGraph.from_arcs(hv["n"], [tuple(a) for a in hv["up"]], [tuple(a) for a in hv["down"]])
    return g, hv


def broom_graph(n_leaf):
    """bucket_ch.rs:5638-5670 as a rank-aligned CCH (apex = n_leaf)."""
    from bfgpu import CchGraph
    n = n_leaf + 1
    wu = (np.arange(n_leaf, dtype=np.uint64) + 1) * 10
    wd = np.arange(n_leaf, dtype=np.uint64) + 1
    up_off = np.minimum(np.arange(n + 1, dtype=np.uint64), n_leaf)
    dn_off = np.zeros(n + 1, dtype=np.uint64)
    dn_off[n] = n_leaf
    g = CchGraph(n=n, up_off=up_off, up_tgt=np.full(n_leaf, n_leaf, dtype=np.uint32), up_w=wu.astype(np.uint32),
                 dn_off=dn_off, dn_tgt=np.arange(n_leaf, dtype=np.uint32), dn_w=wd.astype(np.uint32))
    return g, wu, wd


def broom_expect(wu, wd, s, t):
    s = np.asarray(s, dtype=np.int64)[:, None]
    t = np.asarray(t, dtype=np.int64)[None, :]
    d = (wu[s] + wd[t]).astype(np.uint64)
    return np.where(s == t, 0, d).astype(np.uint32)


def grid_graph(rows, cols=None, chord_frac=0.0, seed=20261203, leaf=8192):
    from bfgpu import GridSpec, cached_grid
    spec = GridSpec(rows=rows, cols=cols or rows, seed=seed, chord_frac=chord_frac, leaf_threshold=leaf)
    return cached_grid(spec, CACHE)


@pytest.fixture(scope="session")
def grid_small():
    return grid_graph(24)


@pytest.fixture(scope="session")
def grid_config1():
    """224x224 jittered grid, the reference's plain CPU configuration."""
    return grid_graph(224)


def progress(config, msg):
    """A progress line that bypasses output capture, so a long GPU step shows
    it is alive while a fixture builds or a test runs."""
    cm = config.pluginmanager.getplugin("capturemanager")
    if cm is None:
        print(f"[progress] {msg}", flush=True)
        return
    with cm.global_and_fixture_disabled():
        print(f"[progress] {msg}", flush=True)


@pytest.fixture(scope="session")
def country_graph(request):
    """The 1414 x 1414 jittered grid with 2 % motorway chords (about 2M NBG
    nodes, an edge-based CCH of about 6.7M nodes): the graph of the matrix,
    isochrone and serving configurations, built once per session (about
    100 s on 16 cores)."""
    import time
    from bfgpu.workload import GRAPH_SEED
    t0 = time.perf_counter()
    progress(request.config, "building the 1414x1414 country graph ...")
    from bfgpu import build_grid
    from bfgpu.workload import grid_spec
    # 7 GB (the edge-based arcs, kept for the customization test, add 0.2 GB): not cached on disk
    g = build_grid(grid_spec(1414, 0.02, GRAPH_SEED), keep_ebg=True)
    progress(request.config, f"country graph n={g.n} m_up={g.m_up} m_down={g.m_down} "
                             f"in {time.perf_counter() - t0:.0f} s")
    return g
