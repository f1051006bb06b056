"""Autotuner for the matrix sweep: label width, lane width K, the DOWN unit
size and the source tile.

The knobs, each applied to the engine before it is timed on the caller's own
workload shape (device-resident, after one untimed run that builds that
shape's DOWN plan):
- label_bits: 16 (packed 16-bit DOWN rows where exact) or 32
  (bfg_ctx_set_label_width);
- lanes: K, the source lanes interleaved per lane group (bfg_ctx_configure);
- plan_rows: the most rows one DOWN unit takes (bfg_ctx_tune);
- source_tile: sources per sweep, all of them or a half or a quarter (the lane
  capacity, bfg_ctx_configure), the role of the reference's source-tile
  picker (route/src/matrix/tile_geometry.rs:203-249).
The search is a coordinate descent: one knob at a time, in the order above,
each left at its fastest value before the next is searched.  Results do not
depend on the tuning (labels are exact), so the tuner only ever trades time.
"""
import time

from .graph import INF

KNOBS = (
    ("label_bits", (16, 32)),
    ("lanes", (64, 32, 16)),
    ("plan_rows", (256, 128, 512)),
    ("source_tile", (1, 2, 4)),  # sources per sweep: all, a half, a quarter
)


def apply(engine, n_src, cfg):
    engine.set_label_width(cfg["label_bits"])
    cap = 0 if cfg["source_tile"] == 1 else (n_src + cfg["source_tile"] - 1) // cfg["source_tile"]
    engine.configure(lanes=cfg["lanes"], lane_capacity=cap)
    engine.tune(plan_rows=cfg["plan_rows"])


def autotune_matrix(engine, d_src, n_src, d_tgt, n_tgt, d_out, reps=2, knobs=KNOBS, log=None):
    """Times compute_matrix_device over the candidates; returns (best
    configuration dict, {candidate label: ms per matrix}) and leaves the best
    applied."""
    from .hip import sync
    cfg = {name: values[0] for name, values in knobs}
    timings = {}

    def run(c):
        key = ",".join(f"{k}={v}" for k, v in c.items())
        if key in timings:
            return timings[key]
        apply(engine, n_src, c)
        engine.compute_matrix_device(d_src, n_src, d_tgt, n_tgt, d_out, INF)  # plan + warm-up
        sync()
        t0 = time.perf_counter()
        for _ in range(reps):
            engine.compute_matrix_device(d_src, n_src, d_tgt, n_tgt, d_out, INF)
        sync()
        timings[key] = 1e3 * (time.perf_counter() - t0) / reps
        if log:
            log(f"autotune {key}: {timings[key]:.1f} ms")
        return timings[key]

    for name, values in knobs:
        best_v, best_t = cfg[name], None
        for v in values:
            c = dict(cfg)
            c[name] = v
            t = run(c)
            if best_t is None or t < best_t:
                best_v, best_t = v, t
        cfg[name] = best_v
    apply(engine, n_src, cfg)
    return cfg, timings
