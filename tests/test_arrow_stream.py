"""Row streaming: matrices as Arrow IPC tiles in the reference's tile schema
(route/src/matrix/arrow_stream.rs:6-15, :102-110), reassembled by a consumer."""
import io

import numpy as np
import pytest

pa = pytest.importorskip("pyarrow")

from bfgpu.arrow_stream import matrix_tile_schema, read_matrix_ipc, stream_matrix_tiles, write_matrix_ipc
from bfgpu.graph import INF
from test_serving import OracleEngine


def test_schema_matches_the_reference():
    s = matrix_tile_schema()
    assert [(f.name, str(f.type), f.nullable) for f in s] == [
        ("src_block_start", "uint32", False), ("dst_block_start", "uint32", False),
        ("src_block_len", "uint16", False), ("dst_block_len", "uint16", False),
        ("durations_ms", "binary", False)]


def test_tiles_cover_the_matrix_once(grid_small):
    eng = OracleEngine(grid_small)
    rng = np.random.default_rng(8)
    s = rng.integers(0, grid_small.n, 37).astype(np.uint32)
    t = rng.integers(0, grid_small.n, 29).astype(np.uint32)
    hit = np.zeros((37, 29), dtype=np.int32)
    for s0, t0, sl, tl, block in stream_matrix_tiles(eng, s, t, chunk_sources=16, tile_sources=5, tile_targets=7):
        assert block.shape == (sl, tl) and sl <= 5 and tl <= 7
        hit[s0:s0 + sl, t0:t0 + tl] += 1
    assert (hit == 1).all()
    assert eng.calls == 3  # one GPU call per chunk of 16 sources


def test_ipc_round_trip_and_scale(grid_small):
    eng = OracleEngine(grid_small)
    rng = np.random.default_rng(9)
    s = rng.integers(0, grid_small.n, 21).astype(np.uint32)
    t = rng.integers(0, grid_small.n, 33).astype(np.uint32)
    want = eng.o.compute_matrix_flat(s, t).reshape(21, 33)
    sink = io.BytesIO()
    n_tiles, n_bytes = write_matrix_ipc(eng, s, t, sink, chunk_sources=8, tile_sources=4, tile_targets=10,
                                        tiles_per_batch=3)
    assert n_bytes == 21 * 33 * 4 and n_tiles == 6 * 4
    assert np.array_equal(read_matrix_ipc(io.BytesIO(sink.getvalue()), 21, 33), want)
    sink = io.BytesIO()
    write_matrix_ipc(eng, s, t, sink, scale=100)
    got = read_matrix_ipc(io.BytesIO(sink.getvalue()), 21, 33)
    exp = np.where(want == INF, INF, np.minimum(want.astype(np.uint64) * 100, INF)).astype(np.uint32)
    assert np.array_equal(got, exp)


@pytest.mark.gpu
def test_ipc_stream_on_gpu(grid_config1):
    from bfgpu import BatchedPhastEngine
    from oracle.oracle import Oracle
    g = grid_config1
    rng = np.random.default_rng(10)
    s = rng.integers(0, g.n, 300).astype(np.uint32)
    t = rng.integers(0, g.n, 250).astype(np.uint32)
    print("[gpu-test] arrow tile stream 300x250", flush=True)
    sink = io.BytesIO()
    with BatchedPhastEngine(g) as e:
        write_matrix_ipc(e, s, t, sink, chunk_sources=128, tile_sources=64, tile_targets=100)
    got = read_matrix_ipc(io.BytesIO(sink.getvalue()), 300, 250)
    assert np.array_equal(got.reshape(-1), Oracle(g).compute_matrix_flat_mt(s, t))


# ---------------------------------------------------------------- Flight matrix rows

class Oracle2chEngine:
    """The oracle behind the engine's two-channel table method (CPU tests)."""

    def __init__(self, g, ul, dl):
        from oracle.oracle import Oracle
        self.o = Oracle(g)
        self.o.attach_length(ul, dl)
        self.calls = 0

    def table_seeded_bounded_routed_2ch(self, src, tgt, threshold=INF):
        self.calls += 1
        return self.o.table_phast_lopsided_2ch(src, tgt, threshold)


def lengths(g, seed):
    rng = np.random.default_rng(seed)
    ul = rng.integers(0, 5000, g.m_up).astype(np.uint32)
    dl = rng.integers(0, 5000, g.m_down).astype(np.uint32)
    ul[g.up_w == INF] = INF
    dl[g.dn_w == INF] = INF
    return ul, dl


def seedsets(g, k, seed):
    rng = np.random.default_rng(seed)
    return [[(int(r), int(p), int(q))] for r, p, q in zip(rng.integers(0, g.n, k), rng.integers(0, 20, k),
                                                           rng.integers(0, 200, k))]


def test_flight_schema_matches_the_reference():
    from bfgpu.arrow_stream import flight_matrix_schema
    assert [(f.name, str(f.type), f.nullable) for f in flight_matrix_schema()] == [
        ("source_idx", "uint32", False), ("target_idx", "uint32", False),
        ("duration_ms", "uint32", False), ("distance_m", "uint32", False)]


def test_flight_batch_cells():
    """flight.rs:532-560: > threshold is INF in both columns, seconds to
    saturating milliseconds, sparse drops INF cells, no lengths -> INF."""
    from bfgpu.arrow_stream import flight_matrix_batch
    t = np.array([[5, INF], [5_000_000, 300]], dtype=np.uint32)
    ln = np.array([[70, 80], [90, 4000]], dtype=np.uint32)
    b = flight_matrix_batch(t, ln, [10, 11], [3, 4], threshold=INF).to_pydict()
    assert b["source_idx"] == [10, 10, 11, 11] and b["target_idx"] == [3, 4, 3, 4]
    assert b["duration_ms"] == [5000, INF, INF, 300000]  # 5e9 ms saturates
    assert b["distance_m"] == [70, INF, 90, 4000]
    b = flight_matrix_batch(t, ln, [10, 11], [3, 4], threshold=200, sparse=True).to_pydict()
    assert b == {"source_idx": [10], "target_idx": [3], "duration_ms": [5000], "distance_m": [70]}
    b = flight_matrix_batch(t, None, [10, 11], [3, 4]).to_pydict()
    assert b["distance_m"] == [INF] * 4
    assert flight_matrix_batch(np.full((1, 2), INF, np.uint32), None, [0], [0, 1], sparse=True) is None


def test_flight_stream_equals_the_table(grid_small):
    from bfgpu.arrow_stream import stream_flight_matrix
    ul, dl = lengths(grid_small, 4)
    eng = Oracle2chEngine(grid_small, ul, dl)
    src, tgt = seedsets(grid_small, 13, 1), seedsets(grid_small, 9, 2)
    et, el, _ = eng.o.table_phast_lopsided_2ch(src, tgt, 600)
    rows = {}
    for b in stream_flight_matrix(eng, src, tgt, threshold=600, tile_sources=5, sparse=True):
        d = b.to_pydict()
        for s, t, ms, m in zip(d["source_idx"], d["target_idx"], d["duration_ms"], d["distance_m"]):
            assert (s, t) not in rows
            rows[(s, t)] = (ms, m)
    assert eng.calls == 3
    exp = {(i, j): (int(et[i * 9 + j]) * 1000, int(el[i * 9 + j]))
           for i in range(13) for j in range(9) if et[i * 9 + j] != INF}
    assert rows == exp


@pytest.mark.gpu
def test_gpu_flight_stream_equals_the_oracle(grid_small):
    from bfgpu import BatchedPhastEngine
    from bfgpu.arrow_stream import stream_flight_matrix
    ul, dl = lengths(grid_small, 5)
    ref = Oracle2chEngine(grid_small, ul, dl)
    src, tgt = seedsets(grid_small, 70, 3), seedsets(grid_small, 40, 4)
    with BatchedPhastEngine(grid_small) as e:
        e.attach_length(ul, dl)
        got = list(stream_flight_matrix(e, src, tgt, threshold=INF, tile_sources=64))
    exp = list(stream_flight_matrix(ref, src, tgt, threshold=INF, tile_sources=64))
    assert len(got) == len(exp) == 2
    for a, b in zip(got, exp):
        assert a.equals(b)


@pytest.mark.parametrize("sparse", [False, True])
def test_flight_ipc_completeness_trailer(grid_small, sparse):
    """flight.rs:3612-3658: a complete stream ends with
    {"complete":true,"total_rows":N,"contract":...}; a producer failure leaves
    no trailer and the reader refuses the stream."""
    from bfgpu.arrow_stream import read_flight_matrix_ipc, write_flight_matrix_ipc
    ul, dl = lengths(grid_small, 6)
    eng = Oracle2chEngine(grid_small, ul, dl)
    src, tgt = seedsets(grid_small, 11, 7), seedsets(grid_small, 8, 8)
    buf = io.BytesIO()
    n = write_flight_matrix_ipc(eng, src, tgt, buf, threshold=500, tile_sources=4, sparse=sparse)
    batches, trailer = read_flight_matrix_ipc(io.BytesIO(buf.getvalue()))
    assert trailer == {"complete": True, "total_rows": n, "contract": "sparse" if sparse else "dense"}
    assert sum(b.num_rows for b in batches) == n
    if not sparse:
        assert n == 11 * 8

    class Failing(Oracle2chEngine):
        def table_seeded_bounded_routed_2ch(self, s, t, threshold=INF):
            if self.calls == 1:
                raise RuntimeError("device lost")
            return super().table_seeded_bounded_routed_2ch(s, t, threshold)

    bad = io.BytesIO()
    with pytest.raises(RuntimeError):
        write_flight_matrix_ipc(Failing(grid_small, ul, dl), src, tgt, bad, tile_sources=4)
    with pytest.raises(ValueError):
        read_flight_matrix_ipc(io.BytesIO(bad.getvalue()))


def test_consumer_closing_early_does_not_hang(grid_small):
    """Closing the tile generator after its first tile (a consumer that stops
    reading) returns at once and leaves no producer blocked on the queue."""
    import threading
    import time
    eng = OracleEngine(grid_small)
    rng = np.random.default_rng(11)
    s = rng.integers(0, grid_small.n, 64).astype(np.uint32)
    t = rng.integers(0, grid_small.n, 16).astype(np.uint32)
    before = threading.active_count()
    gen = stream_matrix_tiles(eng, s, t, chunk_sources=4, tile_sources=4, tile_targets=16)
    next(gen)
    t0 = time.perf_counter()
    gen.close()
    assert time.perf_counter() - t0 < 5.0
    assert threading.active_count() <= before
    assert eng.calls < 16  # the producer stopped instead of computing every chunk


def test_writer_error_stops_the_producer(grid_small):
    """A sink that fails mid-stream surfaces its error and ends the producer."""
    eng = OracleEngine(grid_small)
    rng = np.random.default_rng(12)
    s = rng.integers(0, grid_small.n, 64).astype(np.uint32)
    t = rng.integers(0, grid_small.n, 16).astype(np.uint32)

    class Broken(io.RawIOBase):
        def __init__(self):
            self.n = 0

        def writable(self):
            return True

        def write(self, b):
            self.n += 1
            if self.n > 3:
                raise OSError("sink closed")
            return len(b)

    with pytest.raises(OSError):
        write_matrix_ipc(eng, s, t, Broken(), chunk_sources=4, tile_sources=4, tile_targets=16, tiles_per_batch=1)
    assert eng.calls < 16
