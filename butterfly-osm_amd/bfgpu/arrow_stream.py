"""Row streaming of matrices: Arrow IPC tiles and Flight matrix rows.

The wire format of the reference's matrix tiles (route/src/matrix/arrow_stream.rs:
6-15, schema :102-110): one record per tile of (src_block_start u32,
dst_block_start u32, src_block_len u16, dst_block_len u16, durations_ms
binary), the durations row-major little-endian u32.  A writer sends the schema,
then record batches as tiles complete (the streaming protocol, :17-22).

stream_matrix_tiles computes the matrix in source chunks on the GPU and yields
each chunk's tiles while the next chunk runs: a worker thread keeps the device
busy with chunk i + 1 while the caller packs and sends chunk i.  Values are
the engine's u32 durations times `scale` (saturating, INF kept), so a caller
holding second weights passes scale=1000 for milliseconds.

stream_flight_matrix emits the Flight matrix rows (source_idx, target_idx,
duration_ms, distance_m) from two-channel seeded tables, one record batch per
block of sources; write_flight_matrix_ipc adds the completeness trailer.
"""
import queue
import threading

import numpy as np

from .graph import INF


def matrix_tile_schema():
    import pyarrow as pa
    return pa.schema([
        pa.field("src_block_start", pa.uint32(), nullable=False),
        pa.field("dst_block_start", pa.uint32(), nullable=False),
        pa.field("src_block_len", pa.uint16(), nullable=False),
        pa.field("dst_block_len", pa.uint16(), nullable=False),
        pa.field("durations_ms", pa.binary(), nullable=False),
    ])


def _scaled(block, scale):
    if scale == 1:
        return block
    wide = block.astype(np.uint64) * np.uint64(scale)
    return np.where((block == INF) | (wide >= INF), INF, wide).astype(np.uint32)


def stream_matrix_tiles(engine, sources, targets, chunk_sources=2048, tile_sources=256, tile_targets=4096,
                        threshold=INF, scale=1):
    """Yields (src_start, dst_start, src_len, dst_len, row-major u32 block) tiles
    covering the S x T matrix, source chunk by source chunk, as they complete."""
    if tile_sources > 0xFFFF or tile_targets > 0xFFFF:
        raise ValueError("tile sides are u16 in the wire format")
    src = np.ascontiguousarray(sources, dtype=np.uint32)
    tgt = np.ascontiguousarray(targets, dtype=np.uint32)
    chunks = [(c0, src[c0:c0 + chunk_sources]) for c0 in range(0, src.size, chunk_sources)]
    done = queue.Queue(maxsize=2)  # bounded: the GPU runs at most one chunk ahead of the consumer
    stop = threading.Event()       # set when the consumer goes away (close, break or an error)

    def put(item):
        # a bounded put that gives up once the consumer has stopped, so the
        # producer never blocks on a queue nobody drains
        while not stop.is_set():
            try:
                done.put(item, timeout=0.05)
                return True
            except queue.Full:
                continue
        return False

    def produce():
        try:
            for c0, s in chunks:
                if stop.is_set():
                    return
                m, _ = engine.compute_matrix_block_gated(s, tgt, threshold)
                if not put((c0, m.reshape(s.size, tgt.size))):
                    return
        except Exception as e:  # handed to the consumer
            put(e)
            return
        put(None)

    worker = threading.Thread(target=produce, daemon=True)
    worker.start()
    try:
        while True:
            item = done.get()
            if item is None:
                break
            if isinstance(item, Exception):
                raise item
            c0, m = item
            for r0 in range(0, m.shape[0], tile_sources):
                for t0 in range(0, m.shape[1], tile_targets):
                    block = _scaled(np.ascontiguousarray(m[r0:r0 + tile_sources, t0:t0 + tile_targets]), scale)
                    yield c0 + r0, t0, block.shape[0], block.shape[1], block
    finally:
        stop.set()
        while True:  # free a producer blocked on a full queue
            try:
                done.get_nowait()
            except queue.Empty:
                break
        worker.join()


def tiles_to_record_batch(tiles):
    """arrow_stream.rs:112-142: one record per tile."""
    import pyarrow as pa
    return pa.record_batch([
        pa.array([t[0] for t in tiles], pa.uint32()),
        pa.array([t[1] for t in tiles], pa.uint32()),
        pa.array([t[2] for t in tiles], pa.uint16()),
        pa.array([t[3] for t in tiles], pa.uint16()),
        pa.array([t[4].astype("<u4").tobytes() for t in tiles], pa.binary()),
    ], schema=matrix_tile_schema())


def write_matrix_ipc(engine, sources, targets, sink, tiles_per_batch=8, **kw):
    """ArrowMatrixWriter (arrow_stream.rs:145-200): the schema, then a record
    batch of up to tiles_per_batch tiles as they complete.  Returns
    (tiles written, duration bytes written)."""
    import pyarrow as pa
    n_tiles = n_bytes = 0
    pending = []
    with pa.ipc.new_stream(sink, matrix_tile_schema()) as w:
        for tile in stream_matrix_tiles(engine, sources, targets, **kw):
            pending.append(tile)
            if len(pending) == tiles_per_batch:
                w.write_batch(tiles_to_record_batch(pending))
                n_tiles += len(pending)
                n_bytes += sum(t[4].nbytes for t in pending)
                pending = []
        if pending:
            w.write_batch(tiles_to_record_batch(pending))
            n_tiles += len(pending)
            n_bytes += sum(t[4].nbytes for t in pending)
    return n_tiles, n_bytes


def flight_matrix_schema():
    """The Flight matrix rows (route/src/server/flight.rs:265-272)."""
    import pyarrow as pa
    return pa.schema([
        pa.field("source_idx", pa.uint32(), nullable=False),
        pa.field("target_idx", pa.uint32(), nullable=False),
        pa.field("duration_ms", pa.uint32(), nullable=False),
        pa.field("distance_m", pa.uint32(), nullable=False),
    ])


def flight_matrix_batch(times, lens, src_idx, dst_idx, threshold=INF, sparse=False):
    """build_matrix_tile_batch (flight.rs:505-576) for one tile: times and lens
    are row-major len(src_idx) x len(dst_idx) (lens may be None: no length
    weights).  A cell above threshold is INF in both columns; a finite time
    goes out as milliseconds (saturating); sparse drops INF cells.  Returns a
    record batch, or None when sparse drops every cell."""
    import pyarrow as pa
    src_idx = np.asarray(src_idx, dtype=np.uint32)
    dst_idx = np.asarray(dst_idx, dtype=np.uint32)
    t = np.asarray(times, dtype=np.uint32).reshape(src_idx.size, dst_idx.size)
    t = np.where(t > np.uint32(threshold), np.uint32(INF), t)
    fin = t != INF
    ms = np.where(fin, np.minimum(t.astype(np.uint64) * 1000, INF), INF).astype(np.uint32)
    if lens is None:
        dist = np.full(t.shape, INF, dtype=np.uint32)
    else:
        dist = np.where(fin, np.asarray(lens, dtype=np.uint32).reshape(t.shape), np.uint32(INF))
    si = np.repeat(src_idx, dst_idx.size).reshape(t.shape)
    di = np.tile(dst_idx, src_idx.size).reshape(t.shape)
    keep = fin if sparse else np.ones(t.shape, dtype=bool)
    if not keep.any():
        return None
    return pa.record_batch([pa.array(si[keep], pa.uint32()), pa.array(di[keep], pa.uint32()),
                            pa.array(ms[keep], pa.uint32()), pa.array(dist[keep], pa.uint32())],
                           schema=flight_matrix_schema())


def stream_flight_matrix(engine, src_seedsets, tgt_seedsets, threshold=INF, tile_sources=256, sparse=False):
    """Yields the Flight matrix rows tile by tile: each block of sources runs as
    one two-channel seeded table on the GPU (bfg_table_seeded_2ch; times and
    lengths along them) and goes out as one record batch.  Same-edge
    conflicts are INF here (the reference recomputes them with its bucket
    engine)."""
    T = len(tgt_seedsets)
    dst = np.arange(T, dtype=np.uint32)
    for s0 in range(0, len(src_seedsets), tile_sources):
        block = src_seedsets[s0:s0 + tile_sources]
        t, ln, _ = engine.table_seeded_bounded_routed_2ch(block, tgt_seedsets, threshold)
        b = flight_matrix_batch(t, ln, np.arange(s0, s0 + len(block), dtype=np.uint32), dst, threshold, sparse)
        if b is not None:
            yield b


def write_flight_matrix_ipc(engine, src_seedsets, tgt_seedsets, sink, threshold=INF, tile_sources=256,
                            sparse=False):
    """The Flight matrix rows as an Arrow IPC stream with the completeness
    trailer of completed_flight_stream (flight.rs:3612-3658): once every tile is
    out, one empty record batch whose custom metadata `app_metadata` is
    {"complete":true,"total_rows":N,"contract":"dense"|"sparse"}.  A producer
    that fails writes no trailer (its exception propagates), so a reader can
    tell a truncated stream from a complete one.  Returns the row count."""
    import pyarrow as pa
    schema = flight_matrix_schema()
    rows = 0
    with pa.ipc.new_stream(sink, schema) as w:
        for b in stream_flight_matrix(engine, src_seedsets, tgt_seedsets, threshold, tile_sources, sparse):
            w.write_batch(b)
            rows += b.num_rows
        meta = '{"complete":true,"total_rows":%d,"contract":"%s"}' % (rows, "sparse" if sparse else "dense")
        empty = pa.record_batch([pa.array([], f.type) for f in schema], schema=schema)
        w.write_batch(empty, custom_metadata={"app_metadata": meta})
    return rows


def read_flight_matrix_ipc(source):
    """Reads a stream written by write_flight_matrix_ipc: returns (record
    batches, trailer dict).  Raises if the trailer is missing or its row count
    disagrees with the rows received."""
    import json
    import pyarrow as pa
    batches, trailer = [], None
    reader = pa.ipc.open_stream(source)
    while True:
        try:
            b, meta = reader.read_next_batch_with_custom_metadata()
        except StopIteration:
            break
        if meta is not None and b"app_metadata" in meta:
            trailer = json.loads(meta[b"app_metadata"])
        else:
            batches.append(b)
    if trailer is None or not trailer.get("complete"):
        raise ValueError("flight matrix stream truncated: no completeness trailer")
    if trailer["total_rows"] != sum(b.num_rows for b in batches):
        raise ValueError("flight matrix stream: trailer row count disagrees with the rows received")
    return batches, trailer


def read_matrix_ipc(source, n_src, n_tgt):
    """Reassembles the S x T matrix from a tile stream (the consumer side)."""
    import pyarrow as pa
    out = np.full((n_src, n_tgt), INF, dtype=np.uint32)
    seen = np.zeros((n_src, n_tgt), dtype=bool)
    for batch in pa.ipc.open_stream(source):
        cols = batch.to_pydict()
        for s0, t0, sl, tl, buf in zip(cols["src_block_start"], cols["dst_block_start"], cols["src_block_len"],
                                       cols["dst_block_len"], cols["durations_ms"]):
            out[s0:s0 + sl, t0:t0 + tl] = np.frombuffer(buf, dtype="<u4").reshape(sl, tl)
            seen[s0:s0 + sl, t0:t0 + tl] = True
    if not seen.all():
        raise ValueError("tile stream does not cover the matrix")
    return out
