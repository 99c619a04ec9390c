"""Serving loop: continuous batching of arriving queries over a resident table (config 4).

Requests arrive on their own schedule (Poisson arrivals at a fixed rate in the benchmark).  The
loop keeps the table resident, and runs up to `lanes` batches on the GPU at once, each on its own
lane (a HIP stream with its own buffers).  The lane count is derived from the device memory left
after the table: a lane's footprint (rdb_batch_lane_bytes) times the lane count stays within a
fraction of the free memory.

Whenever a lane is free and requests are waiting, the waiting requests (up to the lane's batch
size) are admitted as one batch.  Requests that arrive while batches are on the GPU wait only for
the next free lane, not for the running batches as a whole.  Finished batches are collected by
polling, without blocking the loop.  A request's latency runs from its arrival to the moment its
batch's results are collected.

Every result is the one the single-query path returns for that query alone (the batched path
evaluates each query's filter and action independently); tests/test_gpu_serving.py checks it.
"""
from __future__ import annotations

import time
from collections import deque

import numpy as np


def poisson_arrivals(n: int, rate: float, seed: int) -> np.ndarray:
    """Arrival times (seconds from 0) of n requests at `rate` per second."""
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate, n))


class Server:
    def __init__(self, engine, table, max_batch: int = 256, row_capacity: int = 65536, memory_fraction: float = 0.5,
                 max_lanes: int = 8, arrays: bool = False):
        self.engine = engine
        self.arrays = arrays  # Mutations results as columns (engine.MutationArrays) instead of row tuples
        self.table = table
        self.max_batch = max_batch
        free, total = engine.device_memory()
        lane = table.lane_bytes(max_batch, row_capacity)
        self.lanes = int(max(1, min(max_lanes, (free * memory_fraction) // max(lane, 1))))
        self.lane_bytes = lane
        self.free_bytes = free
        engine.set_tuning("batch_lanes", self.lanes)
        engine.set_tuning("batch_max_queries", max_batch)
        engine.set_tuning("batch_row_capacity", row_capacity)

    def run(self, requests: list, arrivals: np.ndarray, realtime: bool = True) -> tuple:
        """Serves `requests` (batched-path query dicts) arriving at `arrivals` (seconds).  Returns
        (results, latencies in seconds, wall seconds, batch sizes)."""
        n = len(requests)
        results = [None] * n
        latency = np.zeros(n)
        pending = deque()
        inflight = []  # (batch, request ids)
        sizes = []
        nxt = 0
        done = 0
        t0 = time.perf_counter()
        while done < n:
            now = time.perf_counter() - t0
            if not realtime and not pending and not inflight and nxt < n:
                # nothing to do until the next arrival: in simulated time, jump to it
                t0 -= arrivals[nxt] - now
                now = arrivals[nxt]
            while nxt < n and arrivals[nxt] <= now:
                pending.append(nxt)
                nxt += 1
            progressed = False
            still = []
            for b, ids in inflight:
                if b.poll():
                    out = b.results(self.arrays)
                    t = time.perf_counter() - t0
                    for i, r in zip(ids, out):
                        results[i] = r
                        latency[i] = t - arrivals[i]
                    done += len(ids)
                    b.free()
                    progressed = True
                else:
                    still.append((b, ids))
            inflight = still
            while pending and len(inflight) < self.lanes:
                # one batch = one lane job: its Mutations queries read one column (the oldest
                # waiting one's); the others wait for the next batch
                ids, skipped, col = [], deque(), None
                while pending and len(ids) < self.max_batch:
                    i = pending.popleft()
                    c = requests[i].get("column") if requests[i]["action"] == "mutations" else None
                    if c is not None and col is not None and c != col:
                        skipped.append(i)
                        continue
                    col = col or c
                    ids.append(i)
                skipped.extend(pending)
                pending = skipped
                inflight.append((self.table.batch_submit([requests[i] for i in ids]), ids))
                sizes.append(len(ids))
                progressed = True
            if not progressed:
                time.sleep(20e-6)
        wall = time.perf_counter() - t0
        return results, latency, wall, sizes


def config4_requests(ds, aa_genes: list, n: int, seed: int, lo: float = 1e-6, hi: float = 0.5,
                     min_proportion: float = 0.05) -> list:
    """The serving mix: 60 % Mutations (nucleotide or one of the amino-acid columns), 25 %
    Aggregated by lineage x country, 15 % n-of-m over 20 positions.  Filters are metadata and
    sequence predicates whose expected selectivities are log-uniform over [lo, hi]: a date range of
    the right width narrowed by a lineage or country of matching frequency."""
    import math

    rng = np.random.default_rng(seed)
    cw = ds.country_weight / ds.country_weight.sum()
    lw = ds.lineage_weight / ds.lineage_weight.sum()
    n_days = ds.spec.n_days
    base = ds.spec.date_base
    hot = np.unique(ds.mut_pos)

    def filt(target: float) -> str:
        # pick a lineage or country close to the target, then cut by a date window for the rest
        if rng.random() < 0.5:
            v = int(np.argmin(np.abs(np.log(lw) - math.log(max(target, 1e-9)))))
            leaf, share = f'(lineage lineage "{ds.lineage_names[v]}" 0)', lw[v]
        else:
            v = int(np.argmin(np.abs(np.log(cw) - math.log(max(target, 1e-9)))))
            leaf, share = f'(str_eq country "{ds.country_names[v]}")', cw[v]
        rest = target / share
        if rest >= 0.999:
            return leaf
        days = max(1, int(round(rest * n_days)))
        a = base + int(rng.integers(0, max(n_days - days, 1)))
        return f"(and {leaf} (date_between date {a} {a + days - 1}))"

    out = []
    for _ in range(n):
        target = 10 ** rng.uniform(math.log10(lo), math.log10(hi))
        u = rng.random()
        if u < 0.6:
            col = "main" if rng.random() < 0.5 else aa_genes[int(rng.integers(0, len(aa_genes)))]
            out.append({"filter": filt(target), "action": "mutations", "column": col, "min_proportion": min_proportion})
        elif u < 0.85:
            out.append({"filter": filt(target), "action": "aggregate", "dims": ["lineage", "country"]})
        else:
            pos = rng.choice(hot, size=20, replace=False) if len(hot) >= 20 else rng.integers(0, ds.genome_length, 20)
            leaves = " ".join(f"(has_mut main {int(p)})" for p in pos)
            k = int(rng.integers(1, 4))
            out.append({"filter": f"(and {filt(min(1.0, target * 50))} (nof {k} 0 {leaves}))", "action": "count"})
    return out
