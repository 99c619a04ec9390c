"""Autotuner of the engine's kernel knobs on a resident table (rdb_set_tuning).

It measures, on the GPU and on the table at hand:

* the dense Mutations kernel's shape: wavefronts sharing each tile or not ("mutation_coop"), loads
  in flight per lane ("mutation_unroll"), chunk groups ("mutation_groups") and 16-bit packed
  counters ("mutation_packed"), timed with HIP events on a filter that touches every chunk;
  fewer loads in flight per lane (4) can win: they leave registers for more wavefronts;
* the dense/sparse switch ("sparse_max_rows"): the dense method costs about the same for any filter
  that touches every chunk, the sparse one grows with the filter's rows; both are timed on filters
  of several cardinalities, the sparse time is fitted as a + b * rows, and the switch is put where
  the fitted line meets the dense time.

No knob changes a result; the tuner only picks the fastest settings.
"""
from __future__ import annotations

import math
import time

import numpy as np


def _time_query(table, expr: str, column: str, reps: int) -> float:
    table.mutations_query(expr, column, 0.05)
    t = time.perf_counter()
    for _ in range(reps):
        table.mutations_query(expr, column, 0.05)
    return (time.perf_counter() - t) / reps


def _kernel_us(engine, table, expr: str, column: str, reps: int) -> float:
    engine.profile(True)
    for _ in range(reps):
        table.mutations_query(expr, column, 0.05)
    p = engine.profile_read()
    engine.profile(False)
    return p["mutations_ms"] / max(p["mutations_launches"], 1) * 1e3


def autotune(engine, table, probes: list, column: str = "main", reps: int = 5, log=None) -> dict:
    """`probes`: filter expressions spanning small to large cardinalities (at least one touching
    every chunk).  Returns the chosen knobs and the measurements; the knobs are applied."""
    say = log or (lambda *_: None)
    cards = []
    for e in probes:
        with table.bitmap(e) as b:
            cards.append(b.cardinality)
    order = np.argsort(cards)
    probes = [probes[i] for i in order]
    cards = [cards[i] for i in order]
    broad = probes[-1]
    engine.set_tuning("sparse_max_rows", -1)
    # 1. dense kernel shape: loads in flight, counter packing (work-groups per CU), chunk groups
    shapes = {}
    for coop in (1, 0):
        for packed in (1, 0):
            for unroll in (2, 4, 8):
                for groups in (0, 12, 16, 18, 21, 24, 32):
                    engine.set_tuning("mutation_coop", coop)
                    engine.set_tuning("mutation_packed", packed)
                    engine.set_tuning("mutation_unroll", unroll)
                    engine.set_tuning("mutation_groups", groups)
                    shapes[(coop, packed, unroll, groups)] = _kernel_us(engine, table, broad, column, reps)
    best = min(shapes, key=shapes.get)
    engine.set_tuning("mutation_coop", best[0])
    engine.set_tuning("mutation_packed", best[1])
    engine.set_tuning("mutation_unroll", best[2])
    engine.set_tuning("mutation_groups", best[3])
    say(f"dense kernel shapes (coop, packed, unroll, groups) -> us: {shapes}; chose {best}")
    # 2. dense/sparse switch from per-query wall times
    dense = []
    sparse = []
    for e in probes:
        engine.set_tuning("sparse_max_rows", -1)
        dense.append(_time_query(table, e, column, reps))
        engine.set_tuning("sparse_max_rows", 1 << 40)
        sparse.append(_time_query(table, e, column, reps))
    engine.set_tuning("sparse_max_rows", -1)
    x = np.asarray(cards, dtype=np.float64)
    ys = np.asarray(sparse)
    # least squares a + b * rows over the probes
    A = np.stack([np.ones_like(x), x], axis=1)
    (a, b), *_ = np.linalg.lstsq(A, ys, rcond=None)
    t_dense = float(np.median(dense))
    if b <= 0:
        threshold = int(x.max())
    else:
        threshold = int(max(0.0, (t_dense - a) / b))
    # never above the largest measured filter on which sparse still won
    won = [c for c, d, s in zip(cards, dense, sparse) if s < d]
    if won:
        threshold = int(min(max(threshold, max(won)), 2 * max(won)))
    else:
        threshold = 0
    engine.set_tuning("sparse_max_rows", threshold)
    say(f"dense/sparse per-query seconds: {[(c, round(d * 1e6, 1), round(s * 1e6, 1)) for c, d, s in zip(cards, dense, sparse)]}; "
        f"sparse fit a={a * 1e6:.1f}us b={b * 1e9:.2f}ns/row; switch at {threshold} rows")
    return {"mutation_coop": best[0], "mutation_packed": best[1], "mutation_unroll": best[2], "mutation_groups": best[3],
            "sparse_max_rows": threshold,
            "dense_shapes_us": {f"c{k[0]}p{k[1]}u{k[2]}g{k[3]}": round(v, 2) for k, v in shapes.items()},
            "probe_cards": [int(c) for c in cards],
            "dense_us": [round(d * 1e6, 1) for d in dense], "sparse_us": [round(s * 1e6, 1) for s in sparse]}


def autotune_batch(engine, table, probes: list, column: str = "main", size: int = 64, reps: int = 5, log=None,
                   cards: list | None = None, margin: float = 0.02, workload: list | None = None) -> dict:
    """Chooses the batched dense kernel ("batch_shared_scan": one pass over the index per query, one
    per four queries or one per eight, the last with work-groups of 512 or 1024 threads,
    "batch_octet_threads") by timing a batch of `size` Mutations queries over the probe filters each
    way, then (given the probes' cardinalities) the dense/sparse switch for batches among those
    cardinalities, then the number of batches in flight ("batch_lanes") and the dense kernel again at
    that concurrency, the switch and the job size.  Every knob value tried returns the per-query
    kernel's results (checked on the last batch of each timing, RuntimeError otherwise); a
    candidate replaces the current choice only when it is faster by more than `margin` (the
    run-to-run spread of a pipelined timing).  With `workload` (the Mutations queries the table will
    serve, as batches of `size`) the batches timed are those queries and the switch candidates are
    their filters' cardinalities (each candidate then splits the workload into dense and sparse
    queries differently); otherwise the batches cycle through the probe filters.  The knobs are
    applied."""
    say = log or (lambda *_: None)
    if workload:
        queries = [dict(q) for q in workload[:size]]
        size = len(queries)
        served = []
        for q in queries:
            with table.bitmap(q["filter"]) as b:
                served.append(b.cardinality)
        cards = sorted(set(int(c) for c in served if c > 0) | set(int(c) for c in (cards or [])))
    else:
        queries = [{"filter": probes[i % len(probes)], "action": "mutations", "column": column,
                    "min_proportion": 0.05} for i in range(size)]
    saved = engine.get_tuning("batch_max_queries")
    engine.set_tuning("batch_max_queries", size)
    times = {}
    results = {}
    engine.set_tuning("batch_octet_groups", 0)
    for mode in ((0, 1024), (1, 1024), (2, 1024), (2, 512)):
        engine.set_tuning("batch_shared_scan", mode[0])
        engine.set_tuning("batch_octet_threads", mode[1])
        results[mode] = [r.rows() for r in table.batch(queries, arrays=True)]
        best_t = float("inf")
        for _ in range(reps):  # the fastest of the repetitions: a step's host noise is one-sided
            t = time.perf_counter()
            table.batch(queries, arrays=True)
            best_t = min(best_t, time.perf_counter() - t)
        times[mode] = best_t
    ref_rows = results[(0, 1024)]
    if any(r != ref_rows for r in results.values()):
        raise RuntimeError("shared-scan and per-query batch results differ")
    best = min(times, key=times.get)
    engine.set_tuning("batch_shared_scan", best[0])
    engine.set_tuning("batch_octet_threads", best[1])
    say(f"batch of {size} probe Mutations queries, ms: per-query kernel {times[(0, 1024)] * 1e3:.2f}, "
        f"four per pass {times[(1, 1024)] * 1e3:.2f}, eight per pass {times[(2, 1024)] * 1e3:.2f} "
        f"(512 threads {times[(2, 512)] * 1e3:.2f}); chose batch_shared_scan={best[0]}, threads {best[1]}")
    # the dense/sparse switch inside a batch: a dense query there shares its pass over the index, so
    # the switch that suits single queries is too high; candidates are the probes' own sizes
    if cards is not None:
        single = engine.get_tuning("sparse_max_rows")
        sw = {}
        for thr in sorted({-1, single, *[int(c) for c in cards]}):
            engine.set_tuning("sparse_max_rows", thr)
            got = table.batch(queries, arrays=True)
            if [r.rows() for r in got] != ref_rows:
                raise RuntimeError("batch results depend on the dense/sparse switch")
            best_t = float("inf")
            for _ in range(reps):
                t = time.perf_counter()
                table.batch(queries, arrays=True)
                best_t = min(best_t, time.perf_counter() - t)
            sw[thr] = best_t
        thr = min(sw, key=sw.get)
        engine.set_tuning("sparse_max_rows", thr)
        say(f"batch dense/sparse switch, ms: {[(k, round(v * 1e3, 3)) for k, v in sw.items()]}; chose {thr}")
    def timed(lanes, steps):
        qps, rows = pipelined(table, queries, lanes, steps=steps, reps=3)
        if rows != ref_rows:
            raise RuntimeError("pipelined batch results differ from the per-query kernel's")
        return qps

    def pick(scores: dict, current):
        """The fastest candidate, unless it beats `current` by no more than the margin."""
        best_key = max(scores, key=scores.get)
        if current in scores and scores[best_key] <= scores[current] * (1.0 + margin):
            return current
        return best_key

    # lanes in flight: concurrent batches overlap one batch's latency-bound kernels with another's
    # (candidates whose buffers fit half the free device memory), then the dense scan again at
    # that concurrency
    free, _ = engine.device_memory()
    lane = table.lane_bytes(size, engine.get_tuning("batch_row_capacity"))
    # at most 8: a query's latency grows with the batches ahead of it
    fit = int(max(1, min(8, (free // 2) // max(lane, 1))))
    lanes_qps = {}
    for lanes in sorted({min(n, fit) for n in (1, 2, 4, 6, 8)}):
        engine.set_tuning("batch_lanes", lanes)
        lanes_qps[lanes] = timed(lanes, max(64, 16 * lanes))
    # fewer lanes (less memory, shorter queues) unless more are faster beyond the noise
    best_lanes = min(lanes_qps)
    for lanes in sorted(lanes_qps):
        if lanes_qps[lanes] > lanes_qps[best_lanes] * (1.0 + margin):
            best_lanes = lanes
    engine.set_tuning("batch_lanes", best_lanes)
    # (scan, work-group size, persistent work-groups of the eight-query scan: 0 = one per CU; fewer
    # leave CUs to the other batches in flight)
    mode_qps = {}
    for mode in ((0, 1024, 0), (1, 1024, 0), (2, 1024, 0), (2, 512, 0), (2, 512, 128), (2, 512, 64)):
        engine.set_tuning("batch_shared_scan", mode[0])
        engine.set_tuning("batch_octet_threads", mode[1])
        engine.set_tuning("batch_octet_groups", mode[2])
        mode_qps[mode] = timed(best_lanes, max(64, 16 * best_lanes))
    best = pick(mode_qps, (best[0], best[1], 0))
    engine.set_tuning("batch_shared_scan", best[0])
    engine.set_tuning("batch_octet_threads", best[1])
    engine.set_tuning("batch_octet_groups", best[2])
    say(f"batches in flight, queries/s: {[(k, round(v)) for k, v in lanes_qps.items()]}; chose {best_lanes}; "
        f"dense scans at that concurrency: {[(k, round(v)) for k, v in mode_qps.items()]}; chose {best}")
    steps = max(64, 16 * best_lanes)
    # the dense/sparse switch and the job size (queries per lane job; the batch of `size` is cut into
    # size / job jobs) at that concurrency
    if cards is not None:
        sw_qps = {}
        current = engine.get_tuning("sparse_max_rows")
        for thr in sorted({-1, current, *[int(c) for c in cards if c < 1 << 20]}):
            engine.set_tuning("sparse_max_rows", thr)
            sw_qps[thr] = timed(best_lanes, steps)
        thr = pick(sw_qps, current)
        if thr > 0:
            # refine between the probes' sizes around the best one
            probe_best = thr
            for f in (0.6, 0.8, 1.25, 1.6, 2.0):
                cand = int(thr * f)
                if cand not in sw_qps:
                    engine.set_tuning("sparse_max_rows", cand)
                    sw_qps[cand] = timed(best_lanes, steps)
            thr = pick(sw_qps, probe_best)
        engine.set_tuning("sparse_max_rows", thr)
        say(f"dense/sparse switch at that concurrency, queries/s: {[(k, round(v)) for k, v in sorted(sw_qps.items())]}; chose {thr}")
    job_qps = {}
    for job in sorted({max(1, size // 4), max(1, size // 2), size}):
        engine.set_tuning("batch_max_queries", job)
        job_qps[job] = timed(best_lanes, steps)
    best_job = pick(job_qps, size)
    say(f"queries per lane job, queries/s: {[(k, round(v)) for k, v in job_qps.items()]}; chose {best_job}")
    engine.set_tuning("batch_max_queries", saved)
    out = {"batch_shared_scan": best[0], "batch_octet_threads": best[1], "batch_octet_groups": best[2],
           "batch_lanes": best_lanes,
           "batch_job_queries": best_job, "probe_batch": size,
           "batch_ms": {"per_query": round(times[(0, 1024)] * 1e3, 3), "four": round(times[(1, 1024)] * 1e3, 3),
                        "eight": round(times[(2, 1024)] * 1e3, 3), "eight_512": round(times[(2, 512)] * 1e3, 3)}}
    if cards is not None:
        out["batch_sparse_max_rows"] = engine.get_tuning("sparse_max_rows")
    return out


def pipelined(table, queries: list, in_flight: int, steps: int, reps: int = 2) -> tuple:
    """Queries/s of `steps` batches of `queries` with up to `in_flight` of them submitted ahead of
    the collection of their results (columnar), after one warm-up batch; the best of `reps` runs
    (host noise only ever slows a run down).  Returns (queries/s, the last batch's results as rows)."""
    from collections import deque

    table.batch(queries, arrays=True)
    best = 0.0
    last = None
    for _ in range(reps):
        pend = deque()
        t = time.perf_counter()
        for _ in range(steps):
            pend.append(table.batch_submit(queries))
            if len(pend) == in_flight:
                b = pend.popleft()
                last = b.results(arrays=True)
                b.free()
        while pend:
            b = pend.popleft()
            last = b.results(arrays=True)
            b.free()
        best = max(best, steps * len(queries) / (time.perf_counter() - t))
    return best, [r.rows() if hasattr(r, "rows") else r for r in last]


def probe_filters(ds, n: int = 8) -> list:
    """Lineage and country equality filters whose expected selectivities spread log-uniformly
    over 1e-5 .. 0.5 (the generator's own value frequencies), plus the full table."""
    lw = ds.lineage_weight / ds.lineage_weight.sum()
    out = []
    for s in np.logspace(math.log10(1e-5), math.log10(0.05), n):
        v = int(np.argmin(np.abs(np.log(lw) - math.log(s))))
        out.append(f'(lineage lineage "{ds.lineage_names[v]}" 0)')
    out.append("(true)")
    return list(dict.fromkeys(out))
