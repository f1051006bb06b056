#!/usr/bin/env python3
"""Benchmark: K-lane batched PHAST S x T duration matrix on one MI355X.

Default workload: 10000 x 10000 matrix (100M u32 cells) on the seeded
1414 x 1414 jittered road grid with 2 % long motorway chords (~2M NBG nodes,
edge-based CCH of ~6.6M nodes), sources/targets 80 % uniform + 20 % in five
Gaussian hot-spots (sigma = 2 % of the extent) snapped to the nearest node.
A step is one full matrix with queries and results resident in HBM.

Prints one JSON line with `roofline` (the sweep kernel against the 8 TB/s HBM
peak) and `cpu_baseline` (the oracle's restatement of the reference's
compute_matrix_flat, timed on a bounded sample on this host).

The same matrix also runs through the router's bucket side
(bfg_table_bucket_device: upward phases of the sources on the CCH and of the
targets on the transposed CCH, joined at their common ancestors); `value` is
the side the router takes with both measured (the bucket at 10k x 10k, every
cell checked equal to the PHAST matrix), and `roofline` is that step's join
kernel against the VALU issue bound, with PHAST's DOWN rows kernel beside it
(`roofline_phast`).

Other modes print their own JSON line instead: --serving (continuous batching
against plain per-request calls), --isochrones L (block-gated isochrones with
uneven limits), --table-2ch S (a two-channel seeded table: times and lengths
along them), --customize (the graph's triangle relaxation on the GPU from its
original arc weights, every weight checked against the CPU builder's).  The matrix line runs the autotuner first (label width, K,
the DOWN unit size, the source tile; --no-autotune skips it).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(REPO, "butterfly-osm_amd")
for _p in (REPO, PKG):
    if _p not in sys.path:
        sys.path.insert(0, _p)

from bfgpu.workload import WORKLOADS, isochrone_lanes, make_queries, serving_requests  # noqa: E402

METRIC = "matrix cells/s 10k×10k u32 bit-exact, % HBM roofline; serving p50/p99 vs plain"
HBM_PEAK_GBPS = 8000.0
# VALU issue peak: 256 CUs x 4 SIMDs, one 64-lane wave instruction per 4 cycles at 2.4 GHz
VALU_PEAK_GINSTR = 256 * 4 * 2.4e9 / 4 / 1e9
INF_U32 = 0xFFFFFFFF

def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def copy_probe(nbytes):
    """One device-to-device copy of nbytes by a plain copy kernel (torch's),
    nothing else on the GPU: its FETCH_SIZE and WRITE_SIZE readings per byte
    copied scale the same counters read for the path's kernels."""
    import torch
    a = torch.ones(nbytes // 4, dtype=torch.int32, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    log(f"copied {nbytes} bytes")


def summarize_counters(path):
    import csv
    import collections
    agg = collections.defaultdict(lambda: [0, 0.0])
    with open(path) as f:
        for r in csv.DictReader(f):
            k = (r["Kernel_Name"], r["Counter_Name"])
            agg[k][0] += 1
            agg[k][1] += float(r["Counter_Value"])
    for (name, ctr), (cnt, tot) in sorted(agg.items(), key=lambda x: -x[1][1]):
        print(f"{ctr:12s} dispatches={cnt:6d} mean={tot / cnt:14.1f} total={tot:16.1f}  {name[:110]}")


def counter_means(path, counter, match):
    """Dispatches and mean value per dispatch of `counter` over the kernels
    whose name contains `match` in a rocprofv3 counter_collection.csv."""
    import csv
    vals = [float(r["Counter_Value"]) for r in csv.DictReader(open(path))
            if r["Counter_Name"] == counter and match in r["Kernel_Name"]]
    return len(vals), (sum(vals) / len(vals) if vals else 0.0)


def traffic_summary(fetch_csv, write_csv, cfetch_csv, cwrite_csv, lanes, workload, copy_bytes=1 << 30,
                    kernel=""):
    """FETCH_SIZE and WRITE_SIZE (KiB per dispatch) of the DOWN-rows kernel
    (or of `kernel`), each scaled by what the counter reads per KiB for a
    plain copy of copy_bytes (read copy_bytes, write copy_bytes)."""
    kern = kernel or f"rows16_kernel<{lanes}, false>"
    n, fetch = counter_means(fetch_csv, "FETCH_SIZE", kern)
    _, write = counter_means(write_csv, "WRITE_SIZE", kern)
    _, cf = counter_means(cfetch_csv, "FETCH_SIZE", "copyBuffer")
    _, cw = counter_means(cwrite_csv, "WRITE_SIZE", "copyBuffer")
    fs, ws = copy_bytes / 1024 / cf, copy_bytes / 1024 / cw
    return {"kernel": kern, "workload": workload, "lanes": lanes, "dispatches": n,
            "fetch_kib_per_launch": fetch, "write_kib_per_launch": write, "copy_bytes": copy_bytes,
            "copy_fetch_kib": cf, "copy_write_kib": cw, "fetch_scale": fs, "write_scale": ws,
            "hbm_bytes_per_launch": 1024 * (fetch * fs + write * ws)}


def committed_traffic(workload, lanes):
    """HBM bytes per launch of the dominant kernel from the committed counter
    summary for this workload (FETCH_SIZE + WRITE_SIZE, scaled by a copy
    kernel of known size), or None."""
    path = os.path.join(REPO, "profiles", f"rows_down_{workload}_k{lanes}_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    return {"bytes_per_launch": d["hbm_bytes_per_launch"], "source": os.path.relpath(path, REPO)}


def committed_join_traffic(S, T, bits=32):
    """HBM bytes per launch of the bucket join kernel (the u32 one, or the
    16-bit one for bits=16) for an S x T matrix from the committed counter
    summary (FETCH_SIZE + WRITE_SIZE, scaled by a copy kernel of known size),
    or None."""
    path = os.path.join(REPO, "profiles", f"join{'16' if bits == 16 else ''}_bucket_{S}x{T}_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    return {"bytes_per_launch": d["hbm_bytes_per_launch"], "source": os.path.relpath(path, REPO)}


def cpu_baseline(g, src, tgt, budget_s, gpu_rows):
    """Single-thread oracle (compute_matrix_flat, as the reference runs it) on a
    bounded prefix of the sources; the sample rows are checked against the GPU."""
    from oracle.oracle import Oracle
    o = Oracle(g)
    t0 = time.perf_counter()
    first = o.compute_matrix_flat_soa(src[:8], tgt)
    dt1 = time.perf_counter() - t0
    n_batches = max(1, min(int(budget_s / max(dt1, 1e-3)), (len(src) + 7) // 8))
    rows = [first]
    t0 = time.perf_counter()
    if n_batches > 1:
        rows.append(o.compute_matrix_flat_soa(src[8:8 * n_batches], tgt))
    dt = time.perf_counter() - t0 + dt1
    cpu = np.concatenate(rows).reshape(-1, len(tgt))
    ok = bool(np.array_equal(cpu, gpu_rows[:cpu.shape[0]]))
    cells = cpu.shape[0] * len(tgt)
    return {"value": cells / dt, "unit": "cells/s", "cores": 1, "kind": "port",
            "sample": f"{cpu.shape[0]} sources x {len(tgt)} targets (first sources of the workload), "
                      f"{dt:.1f} s single-thread, orc_compute_matrix_flat_soa (K=8 lanes, "
                      f"batched_phast.rs:756-917 semantics)",
            "bit_exact_vs_gpu_rows": ok}, ok


def cpu_baseline_bucket(g, src, tgt, S, T, gpu_block, k=4):
    """The oracle's bucket many-to-many (table_bucket_parallel semantics,
    single thread) on k sources x k targets: (k + k) sweeps whose per-sweep
    time, times the S + T sweeps of the full matrix, gives its cells/s; the
    k x k cells are checked against the GPU's."""
    from oracle.oracle import Oracle
    o = Oracle(g)
    o.table_bucket(src[:1], tgt[:1])  # builds the flats (untimed)
    t0 = time.perf_counter()
    m = o.table_bucket(src[:k], tgt[:k]).reshape(k, k)
    dt = time.perf_counter() - t0
    per_sweep = dt / (2 * k)
    full_s = per_sweep * (S + T)
    return {"value": S * T / full_s, "unit": "cells/s", "cores": 1, "kind": "port",
            "sample": f"orc_table_bucket on {k} sources x {k} targets ({2 * k} sweeps, {dt:.1f} s single-thread), "
                      f"scaled by sweeps to the {S} x {T} matrix's {S + T} sweeps",
            "bit_exact_vs_gpu_block": bool(np.array_equal(m, gpu_block[:k, :k]))}


def cpu_baseline_mt(g, src, tgt, budget_s):
    from oracle.oracle import Oracle
    o = Oracle(g)
    nth = os.cpu_count() or 1
    nth = min(nth, 16)
    k = 8 * nth
    t0 = time.perf_counter()
    o.compute_matrix_flat_mt(src[:k], tgt, nth)
    dt = time.perf_counter() - t0
    return {"value": k * len(tgt) / dt, "unit": "cells/s", "cores": nth, "kind": "port",
            "sample": f"{k} sources x {len(tgt)} targets, one K=8 batch per thread"}


def bucket_bench(eng, d_src, d_tgt, S, T, d_phast, args, phast_ms, src=None, tgt=None):
    """The router's bucket side on the same device-resident matrix
    (bfg_table_bucket_device: upward phases of the sources on the CCH and of
    the targets on the transposed CCH, joined at their common ancestors),
    timed like the PHAST steps; every cell compared with the PHAST matrix.
    Returns its line and the side the router takes with both measured."""
    from bfgpu.hip import DeviceArray, sync
    from bfgpu.router import MatrixRouter, update_cost_ewma
    t0 = time.perf_counter()
    eng.pair_reverse()
    log(f"transposed context paired in {time.perf_counter() - t0:.1f}s")
    d_b = DeviceArray(S * T, np.uint32)

    def bstep():
        return eng.table_bucket_device(d_src.ptr, S, d_tgt.ptr, T, d_b.ptr)
    for i in range(args.warmup):
        st = bstep()
        log(f"bucket warmup {i}: up {st.upward_time_ms:.1f} ms join {st.downward_time_ms:.1f} ms")
    sync()
    up = join = kern = 0.0
    visits = redone = 0
    bits = 32
    t0 = time.perf_counter()
    for i in range(args.steps):
        st = bstep()
        up += st.upward_time_ms
        join += st.downward_time_ms
        kern += st.join_ms
        visits += st.join_visits
        redone += st.join_tiles_u32
        bits = st.join_label_bits
        log(f"bucket step {i}: up {st.upward_time_ms:.1f} ms join {st.downward_time_ms:.1f} ms")
    sync()
    ms = 1000.0 * (time.perf_counter() - t0) / args.steps
    same = bool(np.array_equal(d_b.to_host(), d_phast.to_host())) if d_phast is not None else None
    log(f"bucket {ms:.1f} ms per matrix, every cell equal to the PHAST matrix: {same}")
    d_b.free()
    r = MatrixRouter(eng)  # the router's choice with both sides measured on this shape
    update_cost_ewma(r.fwd_ns, max(int(phast_ms * 1e6) // S, 1))
    share = join / (up + join) if up + join > 0 else 0.5
    update_cost_ewma(r.up_ns, int(ms * 1e6 * (1 - share)) // (S + T))
    update_cost_ewma(r.cell_ns, max(int(ms * 1e6 * share) // (S * T), 1))
    side = r.side(S, T)
    # the join kernel against the VALU issue bound: a visit is one label
    # position of one 64 x 64 tile, 4096 cell updates.  u32: one v_add_u32
    # (clamp) each and one v_min3_u32 per two (1.5 wave instructions per 64
    # updates), 64 source and 64 target labels (512 bytes) read per visit.
    # 16-bit: one v_pk_add_u16 (clamp) and one v_pk_min_u16 per two updates
    # (1.0 per 64), 256 bytes per visit; the tiles with a cell past the 16-bit
    # cap are recomputed by the u32 kernel inside the same timed span
    kms, vis = kern / args.steps, visits / args.steps
    per64, vbytes = (1.0, 256) if bits == 16 else (1.5, 512)
    instr = vis * 4096 * per64 / 64
    achieved = instr / (kms / 1000.0) / 1e9 if kms > 0 else 0.0
    hbm = (vis * vbytes + 4 * S * T) / (kms / 1000.0) / 1e9 if kms > 0 else 0.0
    # against the HBM roofline the metric names, with the bound that limits the
    # kernel (VALU issue) beside it
    traffic = committed_join_traffic(S, T, bits)
    kname = ("meet_join16_kernel (the bucket join on 16-bit label copies: 64 x 64 min-plus tiles, "
             "packed u16 cell pairs)" if bits == 16 else "meet_join_kernel (the bucket join: 64 x 64 min-plus tiles)")
    roof = {"bound": "hbm", "kernel": kname,
            "achieved": hbm, "peak": HBM_PEAK_GBPS, "unit": "GB/s", "frac": hbm / HBM_PEAK_GBPS,
            "traffic": traffic["bytes_per_launch"] if traffic else None,
            "traffic_source": traffic["source"] if traffic else None,
            "launches": 1, "avg_launch_ms": kms, "visits_per_launch": vis, "cell_updates_per_launch": vis * 4096,
            "algorithmic_bytes_per_launch": vis * vbytes + 4 * S * T,
            "label_bits": bits, "tiles_recomputed_u32_per_launch": redone / args.steps,
            "limited_by": "valu issue",
            "valu": {"achieved": achieved, "peak": VALU_PEAK_GINSTR,
                     "unit": f"G wave-instructions/s (algorithmic: {per64} per 64 cell updates)",
                     "frac": achieved / VALU_PEAK_GINSTR}}
    host = None
    if args.host_steps > 0 and src is not None:
        # the same matrix delivered to pinned host memory (PCIe included)
        from bfgpu.hip import PinnedArray
        pinned = PinnedArray(S * T)
        eng.table_bucket_parallel_bounded(src, tgt, INF_U32, out=pinned.a)  # untimed
        t1 = time.perf_counter()
        for _ in range(args.host_steps):
            eng.table_bucket_parallel_bounded(src, tgt, INF_U32, out=pinned.a)
        hms = 1000.0 * (time.perf_counter() - t1) / args.host_steps
        host = {"ms_per_matrix": hms, "cells_per_s": S * T / (hms / 1000.0), "steps": args.host_steps,
                "entry_point": "bfg_table_bucket", "memory": "pinned host (hipHostMalloc)"}
        log(f"bucket host-delivered {hms:.1f} ms per matrix")
        pinned.free()
    line = {"ms_per_step": ms, "cells_per_s": S * T / (ms / 1000.0), "steps": args.steps, "host_delivered": host,
            "phases_ms_per_step": {"up_both_sides_overlapped": up / args.steps, "join_with_reach_and_pairs":
                                   join / args.steps, "join_kernel": kms},
            "entry_point": "bfg_table_bucket_device", "every_cell_equal_to_phast": same,
            "bottom_pairs": int(st.nodes_swept), "roofline": roof}
    return line, side


def serve_stream(loop_cls, eng, reqs, arrivals):
    """Submits reqs at their arrival offsets (seconds) and returns per-request
    latencies (s), the span from the first arrival to the last completion (s),
    the results and the loop's round count."""
    with loop_cls(eng) as loop:
        t0 = time.perf_counter()
        futs = []
        for r, a in zip(reqs, arrivals):
            wait = t0 + a - time.perf_counter()
            if wait > 0:
                time.sleep(wait)
            futs.append(loop.submit(r))
        res = [f.result() for f in futs]
        rounds = loop.rounds
    lat = np.array([f.completed - f.submitted for f in futs])
    span = max(f.completed for f in futs) - t0
    return lat, span, res, rounds


def serving_bench(eng, g, args, rows, chords):
    from bfgpu.serving import PlainLoop, ServingLoop
    n = args.serving_requests
    t0 = time.perf_counter()
    eng.pair_reverse()  # the bucket side of large matrices, both loops
    log(f"transposed context paired in {time.perf_counter() - t0:.1f}s")
    warm = serving_requests(g, 8, args.seed + 7)
    t0 = time.perf_counter()
    for r in warm:
        serve_stream(PlainLoop, eng, [r], [0.0])
    cal = serving_requests(g, 24, args.seed + 8)
    t0 = time.perf_counter()
    serve_stream(PlainLoop, eng, cal, [0.0] * len(cal))
    plain_rate = len(cal) / (time.perf_counter() - t0)
    log(f"plain capacity ~{plain_rate:.1f} requests/s")
    levels = []
    exact = True
    for k, load in enumerate((0.5, 1.0, 2.0)):
        lam = load * plain_rate
        reqs = serving_requests(g, n, args.seed + 10 + k)
        arrivals = np.cumsum(np.random.default_rng(args.seed + 20 + k).exponential(1.0 / lam, n))
        arrivals -= arrivals[0]
        row = {"offered_rps": lam, "load_vs_plain_capacity": load}
        results = {}
        for name, cls in (("plain", PlainLoop), ("serving", ServingLoop)):
            lat, span, res, rounds = serve_stream(cls, eng, reqs, arrivals)
            results[name] = res
            row[name] = {"p50_ms": 1e3 * float(np.percentile(lat, 50)), "p99_ms": 1e3 * float(np.percentile(lat, 99)),
                         "throughput_rps": n / span, "gpu_rounds": rounds}
            log(f"load {load}: {name} p50 {row[name]['p50_ms']:.1f} ms p99 {row[name]['p99_ms']:.1f} ms "
                f"{row[name]['throughput_rps']:.1f} req/s in {rounds} rounds")
        same = all(np.array_equal(a, b) for a, b in zip(results["plain"], results["serving"]))
        exact &= same
        row["bit_exact_vs_plain"] = same
        levels.append(row)
    top = levels[-1]
    res = {
        "metric": "serving p50/p99 vs plain (mixed matrix + isochrone stream, continuous batching)",
        "value": top["serving"]["throughput_rps"], "unit": "requests/s", "n_gpus": 1,
        "higher_is_better": True, "dtype": "u32", "data": "synthetic",
        "config": {"workload": f"Poisson mixed stream, {n} requests per load: 70% matrices S,T log-uniform in "
                               f"[10, 2000], 30% isochrones with limits 5-60 min; {rows}x{rows} jittered grid "
                               f"+ {int(chords * 100)}% chords, n={g.n}",
                   "loads_vs_plain_capacity": [0.5, 1.0, 2.0], "plain_capacity_rps": plain_rate},
        "levels": levels, "bit_exact": exact,
    }
    print(json.dumps(res), flush=True)


def isochrone_bench(eng, g, args, rows, chords):
    """L isochrones from uniform random origins with uneven limits of 5-60
    minutes in one block-gated call (bfg_isochrone_lists_device): reach bits of
    every node and the (rank, time <= limit) lists of every lane, both left
    in HBM.  The unit is settled node-lanes (list entries) per second."""
    from bfgpu.hip import DeviceArray, sync
    L = args.isochrones
    org, lim = isochrone_lanes(g, L, args.seed + 30)
    words = (g.n + 63) // 64
    d_bits = DeviceArray(L * words, np.uint64)
    # a counting call (capacity 0) sizes the pair buffer
    start, length, total, rc = eng.isochrone_lists_device(org, lim, d_bits.ptr, None, 0)
    log(f"isochrones: {total} settled node-lanes ({total / (L * g.n):.3f} of all), sizing call rc={rc}")
    d_pairs = DeviceArray(2 * max(total, 1), np.uint32)
    eng.isochrone_lists_device(org, lim, d_bits.ptr, d_pairs.ptr, total)  # warm-up
    sync()
    times = []
    for i in range(args.steps):
        t0 = time.perf_counter()
        start, length, tot, rc = eng.isochrone_lists_device(org, lim, d_bits.ptr, d_pairs.ptr, total)
        sync()
        times.append(time.perf_counter() - t0)
        st = eng.last_stats(L)
        log(f"isochrones {i}: {1e3 * times[-1]:.1f} ms wall, up {st.upward_time_ms:.1f} sweep "
            f"{st.downward_time_ms:.1f} out {st.output_time_ms:.1f} ms, batches {st.batches}, "
            f"open segments {st.gate_seg_open}/{st.gate_seg_units}, trees {st.gate_bt_open}/{st.gate_bt_units}")
        if rc != 0 or tot != total:
            raise RuntimeError("isochrone lists changed between identical calls")
    wall = float(np.median(times))
    # bit-exact sample: the first lanes' lists and bits against the oracle's seeded PHAST
    k = 8
    exact = None
    if not args.no_cpu:
        sys.path.insert(0, REPO)
        from oracle.oracle import Oracle
        o = Oracle(g)
        exact = True
        bits = d_bits.to_host(0, k * words).reshape(k, words)
        for i in range(k):
            exp = np.array(o.seeded([(int(org[i]), 0)], int(lim[i])), dtype=np.uint32).reshape(-1, 2)
            got = d_pairs.to_host(2 * int(start[i]), 2 * int(length[i])).reshape(-1, 2)
            exact &= bool(np.array_equal(got, exp))
            inside = np.zeros(words * 64, dtype=bool)
            inside[exp[:, 0]] = True
            exact &= bool(np.array_equal(bits[i], np.packbits(inside, bitorder="little").view(np.uint64)))
        log(f"isochrone sample of {k} lanes bit-exact vs the oracle: {exact}")
    res = {
        "metric": "isochrone settled node-lanes/s (block-gated, uneven limits, one call)",
        "value": total / wall, "unit": "node-lanes/s", "n_gpus": 1, "steps": args.steps,
        "ms_per_step": 1e3 * wall, "higher_is_better": True, "dtype": "u32", "data": "synthetic",
        "config": {"workload": f"{L} isochrones, limits uniform from {{5,10,15,30,45,60}} min, uniform random "
                               f"origins; {rows}x{rows} jittered grid + {int(chords * 100)}% chords, n={g.n}",
                   "lanes": L, "output": "reach bits and (rank, time) lists of every lane, in HBM"},
        "settled_node_lanes": total, "settled_fraction": total / (L * g.n),
        "phases_ms": {"up": st.upward_time_ms, "sweep": st.downward_time_ms, "output": st.output_time_ms,
                      "batches": st.batches},
        "gate": {"segment_units": st.gate_seg_units, "segment_open": st.gate_seg_open,
                 "tree_units": st.gate_bt_units, "tree_open": st.gate_bt_open},
        "bit_exact_sample": exact,
    }
    print(json.dumps(res), flush=True)


def synthetic_lengths(g, seed):
    """Length-along-time weights for the synthetic CCH: 13.9 m per second of
    time (50 km/h) with +-20 % per-arc jitter, INF where the time is INF."""
    rng = np.random.default_rng(seed)
    out = []
    for w in (g.up_w, g.dn_w):
        ln = (w.astype(np.float64) * 13.9 * rng.uniform(0.8, 1.2, w.size)).astype(np.uint64)
        ln = np.minimum(ln, INF_U32 - 1).astype(np.uint32)
        ln[w == INF_U32] = INF_U32
        out.append(ln)
    return out


def table2ch_bench(eng, g, args, rows, chords, src, tgt):
    """An S x S two-channel seeded table (times and lengths along them,
    table_phast_lopsided_2ch semantics): each query point is one or two seeds
    (rank, part_time, part_len), the second on an UP neighbour, as a snapped
    phantom node's two edge ends."""
    S = args.table_2ch
    rng = np.random.default_rng(args.seed + 40)

    def seedsets(pts):
        out = []
        for r in pts[:S]:
            r = int(r)
            sets = [(r, int(rng.integers(0, 20)), int(rng.integers(0, 300)))]
            a, b = int(g.up_off[r]), int(g.up_off[r + 1])
            if b > a and rng.random() < 0.5:
                sets.append((int(g.up_tgt[rng.integers(a, b)]), int(rng.integers(0, 20)), int(rng.integers(0, 300))))
            out.append(sets)
        return out
    ss, ts = seedsets(src), seedsets(tgt)
    ul, dl = synthetic_lengths(g, args.seed + 41)
    t0 = time.perf_counter()
    eng.attach_length(ul, dl)
    log(f"length weights attached in {time.perf_counter() - t0:.1f}s")
    eng.table_seeded_bounded_routed_2ch(ss[:64], ts, INF_U32)  # warm-up
    times = []
    for i in range(args.steps):
        t0 = time.perf_counter()
        tm, lm, conf = eng.table_seeded_bounded_routed_2ch(ss, ts, INF_U32)
        times.append(time.perf_counter() - t0)
        log(f"2ch table {i}: {1e3 * times[-1]:.1f} ms wall")
    wall = float(np.median(times))
    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, REPO)
        from oracle.oracle import Oracle
        o = Oracle(g)
        o.attach_length(ul, dl)
        t0 = time.perf_counter()
        k = 0
        exact = True
        while k < S and time.perf_counter() - t0 < args.cpu_budget:
            et, el, ec = o.table_phast_lopsided_2ch(ss[k:k + 1], ts, INF_U32)
            exact &= bool(np.array_equal(et, tm[k * S:(k + 1) * S]) and np.array_equal(el, lm[k * S:(k + 1) * S])
                          and np.array_equal(ec, conf[k * S:(k + 1) * S]))
            k += 1
        el_s = time.perf_counter() - t0
        cpu = {"value": k * S / el_s, "unit": "cells/s", "cores": 1, "kind": "port",
               "sample": f"oracle table_phast_lopsided_2ch, first {k} of {S} sources x {S} targets, single thread",
               "bit_exact_sample": exact}
        log(f"CPU 2ch table: {k} sources in {el_s:.1f}s, bit-exact vs GPU: {exact}")
    res = {
        "metric": "two-channel seeded table cells/s (time + length along it)",
        "value": S * S / wall, "unit": "cells/s", "n_gpus": 1, "steps": args.steps,
        "ms_per_step": 1e3 * wall, "higher_is_better": True, "dtype": "u32x2", "data": "synthetic",
        "config": {"workload": f"{S}x{S} seeded 2-channel table, 1-2 seeds per point, unbounded; "
                               f"{rows}x{rows} jittered grid + {int(chords * 100)}% chords, n={g.n}; "
                               f"synthetic lengths (13.9 m/s +-20 %)"},
        "conflicts": int(conf.sum()),
        "cpu_baseline": cpu,
    }
    print(json.dumps(res), flush=True)


def customize_bench(g, args, rows, chords):
    """Triangle relaxation (customization) of the whole CCH on the GPU from the
    original arc weights (shortcuts INF), bfg_triangle_relax: one call uploads
    the topology, orders the apexes by dependency level, relaxes every lower
    triangle once and downloads weights and middles.  Every weight is checked
    against the CPU builder's customized graph.  The unit is triangles (a DOWN
    arc x -> m and an UP arc m -> y, both finite) per second; the CPU baseline
    is the oracle's restatement of triangle_relax_parallel on a 160 x 160
    grid (about 20 s), single thread."""
    from bfgpu import original_arc_weights, triangle_relax_parallel
    t0 = time.perf_counter()
    uw, dw, usc, dsc = original_arc_weights(g)
    log(f"original weights in {time.perf_counter() - t0:.1f}s: {int(usc.sum())} UP and {int(dsc.sum())} "
        f"DOWN shortcuts")
    walls, kernels = [], []
    r = None
    for i in range(args.warmup + args.steps):
        t0 = time.perf_counter()
        r = triangle_relax_parallel(g, uw, dw)
        w = time.perf_counter() - t0
        log(f"relax {i}: {1e3 * w:.0f} ms wall, kernels {r.stats.kernel_ms:.1f} ms, {r.stats}")
        if i >= args.warmup:
            walls.append(w)
            kernels.append(r.stats.kernel_ms)
    exact = bool(np.array_equal(r.up_weights, g.up_w) and np.array_equal(r.down_weights, g.dn_w))
    log(f"every weight equal to the builder's: {exact}")
    kms = float(np.median(kernels))
    tri = r.stats.triangles
    cpu = None
    if not args.no_cpu:
        from bfgpu import GridSpec, build_grid
        from oracle import oracle as O
        small = build_grid(GridSpec(rows=160, cols=160, seed=args.seed))
        su, sd, _, _ = original_arc_weights(small)
        rs = triangle_relax_parallel(small, su, sd)
        t0 = time.perf_counter()
        ou, od, om, odm, passes = O.triangle_relax(small, su, sd)
        cw = time.perf_counter() - t0
        same = bool(np.array_equal(ou, rs.up_weights) and np.array_equal(od, rs.down_weights)
                    and np.array_equal(om, rs.up_middles) and np.array_equal(odm, rs.down_middles))
        # the oracle evaluates every triangle once per pass
        cpu = {"value": rs.stats.triangles * passes / cw, "unit": "triangles/s", "cores": 1, "kind": "port",
               "sample": f"orc_triangle_relax (triangle_relax_parallel, customization.rs:1083-1200) on the "
                         f"160x160 grid (n={small.n}, {rs.stats.triangles} triangles), {passes} passes in "
                         f"{cw:.1f} s; weights and middles equal to the GPU's: {same}",
               "bit_exact_vs_gpu": same}
        log(f"cpu baseline: {cpu}")
    # algorithmic bytes: per triangle the arc m -> y (target and word, 12 B) and
    # the arc x -> y located and read (12 B); per lowering candidate an 8-byte
    # atomic min
    algo = 24 * tri + 8 * r.stats.improvements
    gbps = algo / (kms / 1e3) / 1e9
    traffic = None
    tpath = os.path.join(REPO, "profiles", f"relax_{args.workload}_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f)["hbm_bytes_per_launch"] * r.stats.launches
    roof = {"bound": "hbm", "kernel": "relax_level_kernel<false> (one launch per dependency level)",
            "achieved": gbps, "peak": HBM_PEAK_GBPS, "unit": "GB/s", "frac": gbps / HBM_PEAK_GBPS,
            "traffic": traffic, "traffic_note": "HBM bytes per call (all levels), from FETCH_SIZE + WRITE_SIZE"
            if traffic else None, "algorithmic_bytes_per_call": algo, "launches": r.stats.launches,
            "avg_launch_ms": kms / max(r.stats.launches, 1),
            "limited_by": "dependent gathers (arc lookup, word check) per 64-triangle chunk"}
    res = {
        "metric": "customization triangles/s (triangle relaxation of the whole CCH, one call)",
        "value": tri / (kms / 1e3), "unit": "triangles/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": float(np.median(walls)) * 1e3, "kernel_ms": kms,
        "higher_is_better": True, "dtype": "u32", "data": "synthetic",
        "config": {"workload": f"customization of the {rows}x{rows} jittered grid + {int(chords * 100)}% chords, "
                               f"n={g.n}, m_up={g.m_up}, m_down={g.m_down}",
                   "initial_weights": "original arcs their edge-based weight, shortcuts INF"},
        "triangles": tri, "improvements": r.stats.improvements, "levels": r.stats.levels,
        "launches": r.stats.launches, "wall_triangles_per_s": tri / float(np.median(walls)),
        "bit_exact_vs_builder": exact, "roofline": roof, "cpu_baseline": cpu,
        "builder_contract_customize_s": g.meta.get("t_contract_s"),
    }
    print(json.dumps(res), flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="matrix10k", choices=sorted(WORKLOADS))
    ap.add_argument("--lanes", type=int, default=64, help="K lanes per group (16, 32, 64)")
    ap.add_argument("--lane-capacity", type=int, default=0, help="sources per device batch (0 = auto)")
    ap.add_argument("--cpu-budget", type=float, default=15.0, help="seconds of single-thread CPU baseline")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-bucket", action="store_true",
                    help="skip the bucket many-to-many (the router's other side) on the same matrix")
    ap.add_argument("--bucket-only", action="store_true",
                    help="instead of the matrix line: the workload's matrix through the bucket many-to-many alone "
                         "(no PHAST steps, no CPU baseline; for profiling); prints its own JSON line")
    ap.add_argument("--no-autotune", action="store_true",
                    help="skip the autotuner (by default it chooses the label width, K, the DOWN unit size and "
                         "the source tile on this workload before the timed steps)")
    ap.add_argument("--autotune", action="store_true", help="(the default; kept for older command lines)")
    ap.add_argument("--host-steps", type=int, default=3,
                    help="matrices also delivered to pinned host memory through bfg_matrix_bounded after the "
                         "timed device-resident steps (reported as host_delivered; 0 = skip)")
    ap.add_argument("--graph-cache", default=os.environ.get("BFG_GRAPH_CACHE", ""))
    ap.add_argument("--seed", type=int, default=20261203)
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="only run one plain device copy of BYTES (calibrates FETCH_SIZE/WRITE_SIZE)")
    ap.add_argument("--summarize-counters", default="", metavar="CSV",
                    help="print the per-dispatch mean of each counter per kernel of a rocprofv3 "
                         "counter_collection.csv and exit")
    ap.add_argument("--serving", action="store_true",
                    help="instead of the matrix line: Poisson streams of mixed matrix and isochrone requests at "
                         "three loads, served by the continuous-batching loop and as plain per-request calls "
                         "(p50/p99 latency, throughput, bit-exact per request); prints its own JSON line")
    ap.add_argument("--serving-requests", type=int, default=100, help="requests per load level")
    ap.add_argument("--isochrones", type=int, default=0, metavar="L",
                    help="instead of the matrix line: L block-gated isochrone sweeps in one call, limits drawn "
                         "from {5,10,15,30,45,60} min (uneven per lane); prints its own JSON line")
    ap.add_argument("--table-2ch", type=int, default=0, metavar="S",
                    help="instead of the matrix line: an S x S two-channel seeded table (times and lengths "
                         "along them) on the workload's graph; prints its own JSON line")
    ap.add_argument("--customize", action="store_true",
                    help="time the GPU triangle relaxation (customization) of the workload's graph")
    ap.add_argument("--traffic-summary", nargs=4, default=None,
                    metavar=("FETCH_CSV", "WRITE_CSV", "COPY_FETCH_CSV", "COPY_WRITE_CSV"),
                    help="print the DOWN-rows kernel's HBM bytes per launch as JSON: its FETCH_SIZE and "
                         "WRITE_SIZE runs, scaled by the same counters of a --copy-probe run of 1 GiB")
    ap.add_argument("--traffic-kernel", default="",
                    help="the kernel --traffic-summary reports (default: the 16-bit DOWN rows kernel)")
    args = ap.parse_args()
    if args.summarize_counters:
        summarize_counters(args.summarize_counters)
        return
    if args.traffic_summary:
        print(json.dumps(traffic_summary(*args.traffic_summary, lanes=args.lanes, workload=args.workload,
                                         kernel=args.traffic_kernel), indent=1))
        return
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.gpus != 1:
        print(f"bench.py runs on exactly one GPU: --gpus {args.gpus} is not supported by this build "
              f"(the path would shard by sources, one replica of the graph per GPU)", file=sys.stderr)
        sys.exit(2)

    from bfgpu import BatchedPhastEngine, GridSpec, build_grid, cached_grid
    from bfgpu.hip import DeviceArray, sync

    rows, chords, S, T = WORKLOADS[args.workload]
    spec = GridSpec(rows=rows, cols=rows, seed=args.seed, chord_frac=chords)
    t0 = time.perf_counter()
    log(f"building graph {rows}x{rows} chords={chords} ...")
    keep_ebg = args.customize  # the original arc weights come from the edge-based arcs
    g = cached_grid(spec, args.graph_cache, verbose=True, keep_ebg=keep_ebg) if args.graph_cache else \
        build_grid(spec, verbose=True, keep_ebg=keep_ebg)
    log(f"graph n={g.n} m_up={g.m_up} m_down={g.m_down} built in {time.perf_counter() - t0:.1f}s")
    if args.customize:
        customize_bench(g, args, rows, chords)
        return
    src = make_queries(g, S, args.seed + 1)
    tgt = make_queries(g, T, args.seed + 2)

    t0 = time.perf_counter()
    eng = BatchedPhastEngine(g, device=0)
    eng.configure(lanes=args.lanes, lane_capacity=args.lane_capacity)
    info = eng.info()
    log(f"context ready in {time.perf_counter() - t0:.1f}s: {info}")
    if args.serving:
        serving_bench(eng, g, args, rows, chords)
        return
    if args.isochrones:
        isochrone_bench(eng, g, args, rows, chords)
        return
    if args.table_2ch:
        table2ch_bench(eng, g, args, rows, chords, src, tgt)
        return
    if args.bucket_only:
        d_src, d_tgt = DeviceArray.from_host(src), DeviceArray.from_host(tgt)
        line, _ = bucket_bench(eng, d_src, d_tgt, S, T, None, args, 0.0, src, tgt)
        line.update({"metric": "bucket matrix cells/s (profiling run)", "value": line["cells_per_s"],
                     "unit": "cells/s", "n_gpus": 1, "dtype": "u32", "data": "synthetic",
                     "config": {"workload": f"{S}x{T} matrix, {rows}x{rows} jittered grid, n={g.n}"}})
        print(json.dumps(line), flush=True)
        return

    d_src = DeviceArray.from_host(src)
    d_tgt = DeviceArray.from_host(tgt)
    d_out = DeviceArray(S * T, np.uint32)

    def step():
        return eng.compute_matrix_device(d_src.ptr, S, d_tgt.ptr, T, d_out.ptr)

    tuned = None
    if not args.no_autotune:
        from bfgpu.autotune import autotune_matrix
        best, timings = autotune_matrix(eng, d_src.ptr, S, d_tgt.ptr, T, d_out.ptr, log=log)
        tuned = dict(best)
        tuned["candidates_ms"] = {k: round(v, 2) for k, v in timings.items()}
        log(f"autotune chose {best}")

    for i in range(args.warmup):
        st = step()
        log(f"warmup {i}: up {st.upward_time_ms:.1f} ms sweep {st.downward_time_ms:.1f} ms "
            f"out {st.output_time_ms:.1f} ms batches {st.batches}")
    sync()
    sweep_ms = up_ms = out_ms = dom_ms = engine_ms = 0.0
    launches = batches = dom_launches = dom_bytes = 0
    t0 = time.perf_counter()
    for i in range(args.steps):
        st = step()
        sweep_ms += st.downward_time_ms
        up_ms += st.upward_time_ms
        out_ms += st.output_time_ms
        engine_ms += st.total_time_ms
        launches += st.sweep_launches
        batches += st.batches
        dom_ms += st.dom_ms
        dom_launches += st.dom_launches
        dom_bytes += st.dom_bytes
        log(f"step {i}: up {st.upward_time_ms:.1f} ms sweep {st.downward_time_ms:.1f} ms "
            f"out {st.output_time_ms:.1f} ms")
    sync()
    wall = time.perf_counter() - t0
    ms_per_step = 1000.0 * wall / args.steps
    cells = S * T
    value = cells / (ms_per_step / 1000.0)

    info = eng.info()
    KL = info["lanes"]
    cap = info["lane_capacity"]
    groups_per_step = sum((min(cap, S - b) + KL - 1) // KL for b in range(0, S, cap))
    n, m_dn = info["n_nodes"], info["m_down"]
    # the subgraph one lane group's sweep covers (the matrix sweep skips the
    # bottom nodes that are no target's ancestor)
    n_s, m_s = st.nodes_swept, st.arcs_swept
    bytes_per_group = 4 * (n_s + 1) + 8 * m_s + 8 * KL * n_s
    sweep_bytes = groups_per_step * bytes_per_group * args.steps
    sweep_gbps = sweep_bytes / (sweep_ms / 1000.0) / 1e9 if sweep_ms > 0 else 0.0
    # the dominant kernel (DOWN rows of the top region): algorithmic bytes of
    # its launches over their summed duration, both from the library
    achieved = dom_bytes / (dom_ms / 1000.0) / 1e9 if dom_ms > 0 else 0.0
    label_bits = tuned["label_bits"] if tuned else 16
    traffic = committed_traffic(args.workload, KL) if label_bits == 16 else None

    n_check = min(S, 64)
    gpu_rows = d_out.to_host(0, n_check * T).reshape(n_check, T)
    bucket = routed = None
    if not args.no_bucket:
        bucket, routed = bucket_bench(eng, d_src, d_tgt, S, T, d_out, args, ms_per_step, src, tgt)
    host = None
    if args.host_steps > 0:
        # the same matrix delivered to host memory (PCIe included), the way a
        # caller without device buffers gets it
        from bfgpu.hip import PinnedArray
        pinned = PinnedArray(S * T)
        eng.compute_matrix_block_gated(src, tgt, INF_U32, out=pinned.a)  # untimed
        t1 = time.perf_counter()
        for _ in range(args.host_steps):
            eng.compute_matrix_block_gated(src, tgt, INF_U32, out=pinned.a)
        host_ms = 1000.0 * (time.perf_counter() - t1) / args.host_steps
        host = {"ms_per_matrix": host_ms, "cells_per_s": S * T / (host_ms / 1000.0), "steps": args.host_steps,
                "entry_point": "bfg_matrix_bounded", "memory": "pinned host (hipHostMalloc)",
                "rows_equal_device_rows": bool(np.array_equal(pinned.a[:n_check * T], gpu_rows.reshape(-1)))}
        log(f"host-delivered {host_ms:.1f} ms per matrix")
        pinned.free()
    cpu = mt = cpu_b = None
    exact = None
    if not args.no_cpu:
        log("cpu baseline (single thread) ...")
        cpu, exact = cpu_baseline(g, src[:n_check], tgt, args.cpu_budget, gpu_rows)
        log(f"cpu baseline {cpu['value']:.3g} cells/s, bit-exact rows: {exact}")
        mt = cpu_baseline_mt(g, src, tgt, args.cpu_budget)
        log(f"cpu baseline all cores {mt['value']:.3g} cells/s ({mt['cores']} threads)")
        if bucket is not None:
            cpu_b = cpu_baseline_bucket(g, src, tgt, S, T, gpu_rows)
            log(f"cpu bucket baseline {cpu_b['value']:.3g} cells/s (extrapolated), exact: "
                f"{cpu_b['bit_exact_vs_gpu_block']}")

    phast_line = {"ms_per_step": ms_per_step, "cells_per_s": value}
    use_bucket = routed == "bucket" and bucket["every_cell_equal_to_phast"]
    if use_bucket:
        # the matrix call a user makes goes where the router sends it
        ms_per_step, value = bucket["ms_per_step"], bucket["cells_per_s"]
    res = {
        "metric": METRIC, "value": value, "unit": "cells/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_per_step, "routed_to": None, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u32", "data": "synthetic",
        "config": {"workload": f"{S}x{T} duration matrix, {rows}x{rows} jittered grid"
                               f"{' + 2% motorway chords' if chords else ''}, edge-based CCH n={n}",
                   "sources": S, "targets": T, "graph_seed": args.seed, "lanes": KL,
                   "lane_capacity": cap, "n_nodes": n, "m_up": info["m_up"], "m_down": m_dn,
                   "etree_depth": info["etree_depth"], "parallelism": "single GPU", "autotune": tuned},
        "roofline_phast": None,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBPS,
                     "traffic": traffic["bytes_per_launch"] if traffic else None,
                     "kernel": (f"rows16_kernel<{KL}, false> (DOWN rows of the top region, packed 16-bit where exact)"
                                if label_bits == 16 else f"rows_kernel<{KL}, false, false> (DOWN rows of the top region)"),
                     "launches": dom_launches, "avg_launch_ms": dom_ms / max(dom_launches, 1),
                     "bytes_per_launch": dom_bytes / max(dom_launches, 1),
                     "traffic_source": traffic["source"] if traffic else None,
                     "sweep": {"bytes_per_lane_group": bytes_per_group, "nodes_swept": n_s, "arcs_swept": m_s,
                               "lane_groups_per_step": groups_per_step,
                               "launches_per_step": launches // args.steps,
                               "ms_per_step": sweep_ms / args.steps, "achieved_gbps": sweep_gbps}},
        "phases_ms_per_step": {"up": up_ms / args.steps, "sweep": sweep_ms / args.steps,
                               "gather": out_ms / args.steps,
                               "batches_wall": engine_ms / args.steps,
                               "host_staging": phast_line["ms_per_step"] - engine_ms / args.steps},
        "phast": phast_line,
        "bucket": bucket,
        "host_delivered": host,
        "cpu_baseline": cpu,
        "cpu_baseline_all_cores": mt,
        "cpu_baseline_bucket": cpu_b,
        "bit_exact_sample": exact,
    }
    if use_bucket:  # the routed step's dominant kernel; PHAST's DOWN rows beside it
        res["roofline_phast"] = res["roofline"]
        res["roofline"] = bucket["roofline"]
    else:
        res.pop("roofline_phast")
    res["routed_to"] = "bucket" if use_bucket else "phast"
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
