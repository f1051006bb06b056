"""GPU parity at the full sizes of the matrix, isochrone and serving
configurations: the 1414 x 1414 jittered grid with 2 % motorway chords (an
edge-based CCH of about 6.7M nodes, a 13.5k-deep elimination tree), built once
per session.

Every check is bit-exact against the oracle (u32 integer work):
- 1000 x 1000 matrix as one plain call with K = 32 lanes, every cell;
- 10k x 10k matrix delivered to the host, 256 sampled rows (hot-spot sources,
  the lowest- and highest-rank sources);
- the bucket many-to-many (bfg_table_bucket) on both matrices: every cell
  equal to the PHAST matrix (1000 x 1000 also against the oracle);
- 4096 isochrones with uneven 5-60 minute limits in one call: on every lane
  the settled list equals the reach bits (ascending ranks, times within the
  limit, the origin at 0); 32 lanes against the oracle's PHAST rows masked to
  the limit (and the bits of a second call) and 4 lanes against its seeded
  PHAST;
- a mixed serving burst: every result equal to the request issued alone, and
  a sample against the oracle;
- the customization of the whole CCH (triangle relaxation from the original
  arc weights, shortcuts INF): every weight equal to the CPU builder's.
"""
import time

import numpy as np
import pytest

from conftest import progress
from oracle.oracle import INF, Oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def country(country_graph, request):
    from bfgpu import BatchedPhastEngine
    t0 = time.perf_counter()
    e = BatchedPhastEngine(country_graph)
    progress(request.config, f"country context ready in {time.perf_counter() - t0:.0f} s: {e.info()}")
    o = Oracle(country_graph)
    yield country_graph, e, o
    e.close()


def oracle_rows_masked(o, origins, limits):
    """The oracle's unbounded PHAST rows masked to each lane's limit: the
    bounded contract (exact for d <= limit, INF beyond), pinned against the
    oracle's bounded and seeded restatements in the CPU tests."""
    m = o.compute_matrix_flat_mt(origins, np.arange(o.n, dtype=np.uint32)).reshape(len(origins), o.n)
    lim = np.asarray(limits, dtype=np.uint32)[:, None]
    return np.where(m > lim, np.uint32(INF), m)


def bits_of(rows, n):
    words = (n + 63) // 64
    inside = np.zeros((rows.shape[0], words * 64), dtype=bool)
    inside[:, :n] = rows != INF
    return np.packbits(inside, axis=1, bitorder="little").view(np.uint64)


def test_config2_matrix_1000_k32(country, request):
    """1000 x 1000 as a single plain call with K = 32, every cell."""
    g, e, o = country
    from bfgpu.workload import make_queries
    s = make_queries(g, 1000, 20261203 + 1)
    t = make_queries(g, 1000, 20261203 + 2)
    e.configure(lanes=32)
    try:
        got, st = e.compute_matrix_flat(s, t)
    finally:
        e.configure(lanes=64)
    progress(request.config, f"config2 GPU {st.total_time_ms:.1f} ms; oracle ...")
    exp = o.compute_matrix_flat_mt(s, t)
    assert np.array_equal(got, exp)
    assert (got != INF).mean() > 0.9
    # the bucket many-to-many (upward phases on the CCH and its transpose,
    # joined at common ancestors): the same cells
    e.pair_reverse()
    mb, sb = e.table_bucket_parallel(s, t)
    progress(request.config, f"config2 bucket {sb.total_time_ms:.1f} ms")
    assert np.array_equal(mb, exp)


def test_config3_matrix_10k_host_rows(country, request):
    """10k x 10k delivered to host memory; 256 rows against the oracle: the 64
    lowest-rank sources (deep in the elimination tree), the 64 highest-rank
    ones (on the top separators) and 128 hot-spot sources."""
    g, e, o = country
    from bfgpu.workload import make_queries, make_queries_tagged
    s, hot = make_queries_tagged(g, 10000, 20261203 + 1)
    t = make_queries(g, 10000, 20261203 + 2)
    got, st = e.compute_matrix_flat(s, t)
    got = got.reshape(10000, 10000)
    progress(request.config, f"config3 GPU {st.total_time_ms:.1f} ms, {st.batches} batch(es); oracle ...")
    order = np.argsort(s, kind="stable")
    rows = np.unique(np.r_[order[:64], order[-64:], np.nonzero(hot)[0][:128]])
    exp = o.compute_matrix_flat_mt(s[rows], t).reshape(rows.size, 10000)
    assert rows.size >= 250
    assert np.array_equal(got[rows], exp)
    # the bucket many-to-many: every one of the 10^8 cells equals the PHAST matrix
    e.pair_reverse()
    mb, sb = e.table_bucket_parallel(s, t)
    progress(request.config, f"config3 bucket {sb.total_time_ms:.1f} ms, bottom pairs {sb.nodes_swept}")
    assert np.array_equal(mb.reshape(10000, 10000), got)


def test_config4_isochrones_4096(country, request):
    """4096 isochrones, uneven limits of 5-60 minutes, in one call."""
    g, e, o = country
    from bfgpu.workload import isochrone_lanes
    org, lim = isochrone_lanes(g, 4096, 20261203 + 30)
    lists, bits = e.isochrone_lists([[(int(r), 0)] for r in org], lim, want_bits=True)
    st = e.last_stats(4096)
    progress(request.config, f"config4 GPU {st.total_time_ms:.1f} ms, {st.batches} batch(es), "
                             f"{sum(len(x) for x in lists)} settled node-lanes")
    # on every lane: the list's ranks are exactly the set reach bits, ascending,
    # every time within the lane's limit, the origin settled at 0
    for i in range(4096):
        lst = lists[i]
        assert int(np.unpackbits(bits[i].view(np.uint8)).sum()) == len(lst)
        r = lst[:, 0]
        assert np.all(np.diff(r.astype(np.int64)) > 0) and np.all(lst[:, 1] <= lim[i])
        assert np.all((bits[i][r // 64] >> (r % 64).astype(np.uint64)) & np.uint64(1) == 1)
        assert lst[np.searchsorted(r, org[i]), 1] == 0
    # every origin reaches itself at 0
    w, b = org // 64, org % 64
    assert np.all((bits[np.arange(4096), w] >> b.astype(np.uint64)) & np.uint64(1) == 1)
    # 32 lanes (every limit class) against the oracle; their bits must match
    # the 4096-lane call, and the bits must equal times <= limit
    rng = np.random.default_rng(5)
    sample = np.unique(np.r_[rng.choice(4096, 26, replace=False),
                             [int(np.nonzero(lim == m * 60)[0][0]) for m in (5, 10, 15, 30, 45, 60)]])
    sb, stimes = e.isochrones(org[sample], lim[sample], want_times=True)
    assert np.array_equal(sb, bits[sample])
    exp = oracle_rows_masked(o, org[sample], lim[sample])
    assert np.array_equal(stimes, exp)
    assert np.array_equal(sb, bits_of(stimes, g.n))
    # the seeded restatement itself (run_phast_bounded_fast_seeded) on 4 lanes
    for i in sample[:4]:
        assert lists[i].tolist() == [list(x) for x in o.seeded([(int(org[i]), 0)], int(lim[i]))]
    # the sampled lanes' lists against the oracle's masked rows
    for j, i in enumerate(sample):
        nz = np.nonzero(exp[j] != INF)[0]
        assert np.array_equal(lists[i][:, 0], nz) and np.array_equal(lists[i][:, 1], exp[j][nz])


def test_config5_serving_burst(country, request):
    """A mixed burst through the serving loop: each result equals the same
    request issued as its own plain call; matrices and isochrone lists
    sampled against the oracle."""
    g, e, o = country
    from bfgpu.serving import IsochroneRequest, MatrixRequest, ServingLoop, run_plain
    from bfgpu.workload import serving_requests
    reqs = serving_requests(g, 40, 20261203 + 50, s_range=(10, 600))
    t0 = time.perf_counter()
    with ServingLoop(e) as loop:
        got = [f.result(timeout=300) for f in [loop.submit(r) for r in reqs]]
    progress(request.config, f"config5 burst of {len(reqs)} in {time.perf_counter() - t0:.1f} s; alone ...")
    for r, res in zip(reqs, got):
        assert np.array_equal(res, run_plain(e, r))
    mats = [k for k, r in enumerate(reqs) if isinstance(r, MatrixRequest)][:4]
    isos = [k for k, r in enumerate(reqs) if isinstance(r, IsochroneRequest)][:3]
    for k in mats:
        assert np.array_equal(got[k], o.compute_matrix_flat_mt(reqs[k].sources, reqs[k].targets))
    for k in isos:
        exp = o.seeded([(reqs[k].origin, 0)], reqs[k].limit)
        assert got[k].tolist() == [list(x) for x in exp]


def test_customize_country_graph(country_graph, request):
    """GPU triangle relaxation of the whole 6.7M-node CCH from its original arc
    weights: every UP and DOWN weight equal to the CPU builder's customized
    weights (the same fixpoint computed by its fused contraction), and a
    sample of elected middles splitting their arcs into two that add up."""
    from bfgpu import customize
    g = country_graph
    t0 = time.perf_counter()
    out, r = customize(g)
    progress(request.config, f"customization in {time.perf_counter() - t0:.1f} s: {r.stats}")
    assert np.array_equal(out.up_w, g.up_w) and np.array_equal(out.dn_w, g.dn_w)
    rng = np.random.default_rng(3)
    relaxed = np.flatnonzero(r.up_middles != INF)
    for i in rng.choice(relaxed, 500, replace=False):
        x = int(np.searchsorted(g.up_off, i, side="right") - 1)
        y, m = int(g.up_tgt[i]), int(r.up_middles[i])
        a = int(g.dn_off[x]) + int(np.searchsorted(g.dn_tgt[g.dn_off[x]:g.dn_off[x + 1]], m))
        b = int(g.up_off[m]) + int(np.searchsorted(g.up_tgt[g.up_off[m]:g.up_off[m + 1]], y))
        assert g.dn_tgt[a] == m and g.up_tgt[b] == y
        assert min(int(g.dn_w[a]) + int(g.up_w[b]), INF) == g.up_w[i]
