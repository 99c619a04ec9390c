"""Training and evaluation of the learned planner (the config-3 workload).

* Trees: seeded filter trees of the batched workload (workloads.TreeGen, depth 1-6).
* Samples: one per operator node of every compiled tree, features from the engine
  (rdb_plan_samples) with the children's exact values, target = normalised log of the node's exact
  cardinality, evaluated by the engine on the GPU.
* Training: the estimator (3 x 128 hidden units, BF16 parameters, FP32 accumulation) on the GPU by
  the engine's own forward / backward / Adam kernels.
* Evaluation on held-out trees: q-error = max(est / exact, exact / est) (with +1 on both) of the
  estimates the planner would use (bottom-up, each node from its children's estimates), for the
  learned estimator and for the independence estimate it replaces; then the batched workload with
  the learned plan against the reference plan: throughput, and every result compared.
"""
from __future__ import annotations

import time

import numpy as np

from . import workloads
from .engine import plan_estimates, plan_samples


def trees(ds, n: int, seed: int) -> list:
    gen = workloads.TreeGen(ds, seed)
    return [gen.tree(int(gen.rng.integers(1, 7))) for _ in range(n)]


def samples(et, filters: list, chunk: int = 2000, log=None):
    xs, ys = [], []
    t0 = time.time()
    for i in range(0, len(filters), chunk):
        f, y, _ = plan_samples(et, filters[i:i + chunk])
        xs.append(f)
        ys.append(y)
        if log:
            log(f"samples: {i + len(filters[i:i + chunk])} trees, {sum(len(a) for a in ys)} nodes, {time.time() - t0:.1f}s")
    return np.concatenate(xs), np.concatenate(ys)


def qerrors(est, exact):
    e = np.asarray(est, np.float64) + 1.0
    x = np.asarray(exact, np.float64) + 1.0
    return np.maximum(e / x, x / e)


def evaluate(et, mlp, filters: list, chunk: int = 2000) -> dict:
    learned, indep, exact, roots = [], [], [], []
    for i in range(0, len(filters), chunk):
        est, ind, ex, tr = plan_estimates(et, filters[i:i + chunk], mlp)
        learned.append(est)
        indep.append(ind)
        exact.append(ex)
        last = np.r_[np.flatnonzero(np.diff(tr.astype(np.int64)) != 0), len(tr) - 1]
        roots.append(np.zeros(len(tr), bool))
        roots[-1][last] = True
    learned, indep, exact, roots = map(np.concatenate, (learned, indep, exact, roots))
    out = {}
    for name, est in (("learned", learned), ("independence", indep)):
        q = qerrors(est, exact)
        out[name] = {"nodes_median": float(np.median(q)), "nodes_p95": float(np.percentile(q, 95)),
                     "roots_median": float(np.median(q[roots])), "roots_p95": float(np.percentile(q[roots], 95))}
    out["nodes"] = int(len(exact))
    out["trees"] = len(filters)
    return out


def speedup(engine, et, queries: list, mlp, steps: int = 3, rounds: int = 5) -> dict:
    """Batched workload under the reference plan, the independence plan and the learned plan
    (results as columns, so that the device work dominates the clock), the modes taking turns for
    `rounds` rounds: `steps` batches one at a time (the host's planning in series with the device)
    and `steps` batches pipelined over the lanes (one batch's planning beside the others' device
    work); the median round per mode; results compared."""
    from collections import deque

    from .engine import MutationArrays

    modes = ("reference", "independence", "learned")
    in_flight = max(1, engine.get_tuning("batch_lanes"))
    sync = {m: [] for m in modes}
    pipe = {m: [] for m in modes}
    outs = {}
    for m in modes:  # warm-up (graph capture, lanes)
        engine.set_planner(m, mlp if m == "learned" else None)
        et.batch(queries, arrays=True)
    for _ in range(rounds):
        for m in modes:
            engine.set_planner(m, mlp if m == "learned" else None)
            t = time.perf_counter()
            for _ in range(steps):
                got = et.batch(queries, arrays=True)
            sync[m].append(time.perf_counter() - t)
            outs[m] = [r.rows() if isinstance(r, MutationArrays) else r for r in got]
            pend = deque()
            t = time.perf_counter()
            for _ in range(steps):
                pend.append(et.batch_submit(queries))
                if len(pend) == in_flight:
                    b = pend.popleft()
                    b.results(arrays=True)
                    b.free()
            while pend:
                b = pend.popleft()
                b.results(arrays=True)
                b.free()
            pipe[m].append(time.perf_counter() - t)
    engine.set_planner("reference")
    qs = {m: len(queries) * steps / float(np.median(sync[m])) for m in modes}
    qp = {m: len(queries) * steps / float(np.median(pipe[m])) for m in modes}
    return {"queries_per_s": {k: round(v, 2) for k, v in qs.items()},
            "learned_vs_reference": round(qs["learned"] / qs["reference"], 4),
            "independence_vs_reference": round(qs["independence"] / qs["reference"], 4),
            "pipelined_queries_per_s": {k: round(v, 2) for k, v in qp.items()},
            "pipelined_learned_vs_reference": round(qp["learned"] / qp["reference"], 4),
            "pipelined_independence_vs_reference": round(qp["independence"] / qp["reference"], 4),
            "in_flight": in_flight,
            "results_identical": outs["learned"] == outs["reference"] == outs["independence"]}
