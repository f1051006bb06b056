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
