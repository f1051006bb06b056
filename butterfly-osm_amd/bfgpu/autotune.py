"""Autotuner for the matrix sweep: lane width K and the DOWN unit size.

Each candidate (K, plan_rows) is applied to the engine (configure, tune) and
timed on the caller's own workload shape, device-resident, after one untimed
run that builds the DOWN plan for that shape.  The fastest candidate is left
applied.  Results do not depend on the tuning (labels are exact), so the
tuner only ever trades time.
"""
import itertools
import time

from .graph import INF


def autotune_matrix(engine, d_src, n_src, d_tgt, n_tgt, d_out, lanes=(64, 32), plan_rows=(128, 256, 512),
                    reps=2, log=None):
    """Times compute_matrix_device over the candidates; returns
    (best (lanes, plan_rows), {candidate: ms per matrix})."""
    from .hip import sync
    timings = {}
    for k, r in itertools.product(lanes, plan_rows):
        engine.configure(lanes=k)
        engine.tune(plan_rows=r)
        engine.compute_matrix_device(d_src, n_src, d_tgt, n_tgt, d_out, INF)  # plan + warm-up
        sync()
        t0 = time.perf_counter()
        for _ in range(reps):
            engine.compute_matrix_device(d_src, n_src, d_tgt, n_tgt, d_out, INF)
        sync()
        timings[(k, r)] = 1e3 * (time.perf_counter() - t0) / reps
        if log:
            log(f"autotune K={k} plan_rows={r}: {timings[(k, r)]:.1f} ms")
    best = min(timings, key=timings.get)
    engine.configure(lanes=best[0])
    engine.tune(plan_rows=best[1])
    return best, timings
