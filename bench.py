#!/usr/bin/env python3
"""Mutations-query benchmark on one MI355X.

A step is one pass of the configuration's query workload over a table resident in HBM:

* config 1 (default): 64 Mutations queries (minProportion 0.05) over 4 M sequences x 29,903 nt,
  filters on date range, country or lineage with selectivities 1e-4..0.5;
* config 2: 1,024 heterogeneous filter trees (depth 1-6 over And/Or/Not/NOf with sequence and
  metadata leaves, exact selectivities log-uniform over 1e-6..0.5), alternately Mutations and
  Aggregated by country, each also returning its cardinality;
* config 3: the config-2 workload with the learned planner: the estimator is trained on the
  engine's exact node cardinalities of seeded trees (--planner-trees) and judged on held-out ones
  (--planner-eval) by q-error against the independence estimate; the batched path then runs with
  the learned plan, and its throughput and results are compared with the reference plan's;
* config 4: serving: 16 M sequences plus ten amino-acid columns resident, requests of the mix
  (60 % Mutations, 25 % Aggregated by lineage x country, 15 % n-of-m over 20 positions) arriving
  as a Poisson process, served by continuous batching; p50/p99 latency and throughput;
* config 0: the 100 k-row CPU-runnable case.

Two paths run the same workload:

* plain: one C-ABI call per query (filter kernel + action kernels + one host round trip);
* batched: the whole step as one batch (rdb_batch_*): filters of all queries in one launch,
  dense or sparse Mutations chosen per query on the device, HIP-graph replay, with the kernel
  knobs picked by the autotuner beforehand (outside the timed region); the next step's batch is
  submitted before the previous one's results are collected, up to one batch per lane (the
  autotuner picks 1-8 lanes), so that concurrent batches share the GPU.

`value` is the batched path's queries/s; the plain path's figures sit beside it.  Every query's
batched result is compared with its plain result.  The CPU baseline is the oracle (a
single-threaded restatement of the reference's CPU operators) on a bounded sample of the same
queries, whose results are also compared with the engine's bit for bit.  Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "lapis-silo_amd"))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "Mutations-query queries/s + p50/p99 latency @4M seqs, bit-exact; GPU vs CPU"


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def percentile(xs, q):
    s = sorted(xs)
    k = min(len(s) - 1, max(0, int(round(q / 100.0 * (len(s) - 1)))))
    return s[k]


def copy_calibration(mib):
    """A plain copy of known size: FETCH_SIZE / WRITE_SIZE of this kernel divided by mib * 1024 gives
    the factor by which the counters' KiB relate to bytes moved."""
    import torch

    n = mib << 20
    src = torch.ones(n, dtype=torch.uint8, device="cuda:0")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    log(f"copied {mib} MiB five times")


def cpu_model() -> str:
    """The host CPU's model name with its family/model/stepping, clock and vector extensions (a
    virtualised host may report only a brand name)."""
    info = {}
    try:
        for line in open("/proc/cpuinfo"):
            if ":" in line:
                k, v = line.split(":", 1)
                info.setdefault(k.strip(), v.strip())
    except OSError:
        return "unknown"
    name = info.get("model name", "unknown")
    flags = set(info.get("flags", "").split())
    vec = [f for f in ("avx512f", "avx512vpopcntdq", "avx2") if f in flags]
    detail = f"family {info.get('cpu family', '?')} model {info.get('model', '?')} stepping {info.get('stepping', '?')}"
    if info.get("cpu MHz"):
        detail += f", {float(info['cpu MHz']) / 1000:.2f} GHz"
    if vec:
        detail += ", " + " ".join(vec)
    return f"{name} ({detail})"


def oracle_query(ot, q):
    if q["action"] == "mutations":
        return ot.mutations(q["filter"], q.get("column", "main"), q["min_proportion"])
    if q["action"] == "aggregate":
        return ot.aggregate(q["filter"], q["dims"])
    return ot.count(q["filter"])


def cpu_legs(spec, queries, want, seconds_one: float, seconds_all: float, threads: int) -> tuple:
    """The oracle (a single-threaded restatement of the reference's CPU operators) timed on a
    bounded sample of the workload: one thread, one query at a time (the reference evaluates a
    query on one worker thread), with per-query latencies; then `threads` threads running
    independent queries of the same workload concurrently on the same table (the reference's worker
    pool of hardware_concurrency threads).  Every oracle result is compared with the engine's
    result for that query (`want`).  Returns (baseline dict, parity dict)."""
    import threading
    from concurrent.futures import ThreadPoolExecutor

    from oracle.oracle import OracleTable

    t0 = time.time()
    ot = OracleTable(spec)
    build_s = time.time() - t0
    log(f"oracle built in {build_s:.1f}s")
    n = len(queries)
    lat = []
    identical = True
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < seconds_one and len(lat) < n:
        i = (len(lat) * 7) % n
        t = time.perf_counter()
        identical &= oracle_query(ot, queries[i]) == want[i]
        lat.append(time.perf_counter() - t)
    one_s = time.perf_counter() - t0
    one = {"queries_per_s": round(len(lat) / one_s, 4), "queries": len(lat),
           "p50_ms": round(percentile(lat, 50) * 1e3, 2), "p99_ms": round(percentile(lat, 99) * 1e3, 2)}
    log(f"cpu baseline, one thread: {one}")
    lock = threading.Lock()
    stop = time.perf_counter() + seconds_all
    results = {"done": 0, "identical": True}

    def worker(k):
        done, same, i = 0, True, k
        while time.perf_counter() < stop:
            j = (i * 7) % n
            same &= oracle_query(ot, queries[j]) == want[j]
            done += 1
            i += threads
        with lock:
            results["done"] += done
            results["identical"] &= same

    t0 = time.perf_counter()
    with ThreadPoolExecutor(threads) as pool:
        list(pool.map(worker, range(threads)))
    all_s = time.perf_counter() - t0
    ot.close()
    identical &= results["identical"]
    cpu = {"value": round(results["done"] / all_s, 4), "unit": "queries/s", "cores": threads, "kind": "port",
           "sample": (f"{results['done']} queries of the {n}-query workload on {threads} threads in {all_s:.1f}s "
                      f"(independent queries, round-robin), oracle build {build_s:.1f}s excluded"),
           "cpu_model": cpu_model(), "host_cpus": os.cpu_count(), "single_thread": one}
    parity = {"queries_compared": len(lat) + results["done"], "identical": bool(identical)}
    log(f"cpu baseline, {threads} threads: {cpu['value']} queries/s; parity {parity}")
    return cpu, parity


def run_plain_query(et, q):
    if q.get("action", "mutations") == "mutations":
        return et.mutations_query(q["filter"], q.get("column", "main"), q["min_proportion"], arrays=True)
    if q["action"] == "aggregate":
        with et.bitmap(q["filter"]) as b:
            return et.aggregate(b, q["dims"])
    return et.count(q["filter"])


def as_rows(results: list) -> list:
    """Mutations results delivered as columns (timed) turned into row tuples (for comparisons)."""
    from silo_amd.engine import MutationArrays

    return [r.rows() if isinstance(r, MutationArrays) else r for r in results]


def scan_roofline(eng, et, queries, prof, n_chunks, log):
    """Roofline object of the batched path's dense counting kernel: algorithmic bytes of its passes
    over the index (the index bytes of the chunks the pass's filters touch, counted once per pass,
    plus one 8 KiB filter tile per query and chunk) over its average launch time.  Every dense
    filter of this workload touches every chunk (checked), so a pass of k queries reads the bytes of
    the dense method under the all-rows filter plus (k - 1) more filter tiles per chunk."""
    launches = prof["scan_launches"]
    if not launches:
        return None
    dense = prof["scan_dense_queries"] / launches
    switch = eng.get_tuning("sparse_max_rows")
    for q in queries:
        if q["action"] != "mutations":
            continue
        with et.bitmap(q["filter"]) as b:
            if b.cardinality > 0 and (switch < 0 or b.cardinality > switch):
                if len(set((b.rows() >> 16).tolist())) != n_chunks:
                    log("scan roofline skipped: a dense filter misses a chunk")
                    return None
    mode = eng.get_tuning("batch_shared_scan")
    per_pass = {0: 1, 1: 4, 2: 8}[mode]
    saved = eng.get_tuning("sparse_max_rows")
    eng.set_tuning("sparse_max_rows", -1)
    with et.bitmap("(true)") as b:
        all_rows = et.mutations_traffic(b, "main")
    eng.set_tuning("sparse_max_rows", saved)
    full_passes, rest = divmod(int(round(dense)), per_pass)
    bytes_per_launch = full_passes * (all_rows + (per_pass - 1) * 8192 * n_chunks)
    if rest:
        bytes_per_launch += all_rows + (rest - 1) * 8192 * n_chunks
    avg_ms = prof["scan_ms"] / launches
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
    kernel = {0: "mutationsCountBatchKernel", 1: "mutationsCountSharedKernel", 2: "mutationsCountOctetKernel"}[mode]
    out = {"bound": "hbm", "kernel": kernel, "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": scan_pmc_bytes(kernel),
           "algorithmic_bytes_per_launch": int(bytes_per_launch), "dense_queries_per_launch": round(dense, 2),
           "queries_per_pass": per_pass, "launches": launches, "avg_launch_ms": round(avg_ms, 5)}
    stream = eng.read_bandwidth(bytes_per_launch, reps=10)
    out["stream_read_GBps_same_bytes"] = round(stream, 1)
    out["frac_of_stream_read_same_bytes"] = round(achieved / stream, 4)
    log(f"batched dense scan roofline: {out}")
    return out


def serve(args):
    """Config 4: 16 M sequences + 10 amino-acid columns resident; Poisson arrivals of the serving
    mix; continuous batching; p50/p99 latency and throughput, results checked against the same
    queries run alone."""
    from silo_amd import datagen, serving, table
    from silo_amd.engine import Engine

    t0 = time.time()
    ds = datagen.generate_config(4, n_rows=args.rows, keep_handle=True)
    layout = datagen.gene_layout(ds.genome_length)
    aa = datagen.translate(ds, layout)
    spec = table.spec_from_dataset(ds, aa_columns=aa)
    log(f"generated {ds.n_rows} rows x {ds.genome_length} nt + {len(aa)} amino-acid columns "
        f"({sum(n for _, _, n in layout)} residues) in {time.time() - t0:.1f}s")
    eng = Engine(0)
    t0 = time.time()
    et = eng.table(spec)
    upload_s = time.time() - t0
    log(f"index built and uploaded in {upload_s:.1f}s: {et.stats()}")
    genes = [g for g, _, _ in layout]
    reqs = serving.config4_requests(ds, genes, n=args.requests, seed=41)
    arrivals = serving.poisson_arrivals(len(reqs), rate=args.rate, seed=42)
    # the single-query dense/sparse switch fitted on this table (serving batches hold few queries)
    from silo_amd import autotune

    tuned = autotune.autotune(eng, et, autotune.probe_filters(ds), log=log)
    log(f"serving switch: sparse method up to {tuned['sparse_max_rows']} rows")
    eng.set_tuning("batch_graphs", args.serving_graphs)
    srv = serving.Server(eng, et, max_batch=256, row_capacity=65536, arrays=True)
    log(f"serving with {srv.lanes} lanes of {srv.lane_bytes / 2**20:.0f} MiB ({srv.free_bytes / 2**30:.1f} GiB free)")
    warm = serving.config4_requests(ds, genes, n=256, seed=40)
    srv.run(warm, serving.poisson_arrivals(len(warm), rate=args.rate, seed=39))
    results, latency, wall, sizes = srv.run(reqs, arrivals)
    results = as_rows(results)
    qps = len(reqs) / wall
    log(f"served {len(reqs)} requests in {wall:.2f}s: {qps:.1f}/s, p50 {np.percentile(latency, 50) * 1e3:.2f} ms, "
        f"p99 {np.percentile(latency, 99) * 1e3:.2f} ms, {len(sizes)} batches (mean {np.mean(sizes):.1f})")
    # the same queries alone (single-query path): a bounded sample, bit for bit
    sample = list(range(0, len(reqs), max(1, len(reqs) // 200)))
    t0 = time.perf_counter()
    identical = all(as_rows([run_plain_query(et, reqs[i])])[0] == results[i] for i in sample)
    plain_qps = len(sample) / (time.perf_counter() - t0)
    log(f"alone == in the mix for {len(sample)} sampled requests: {identical}; single-query path {plain_qps:.1f}/s")
    # capacity: the offered rate raised until p99 passes the target or throughput falls behind
    sweep = []
    capacity = None
    for rate in [float(r) for r in args.sweep_rates.split(",") if r.strip()]:
        n_req = int(max(2000, rate * args.sweep_seconds))
        sreqs = serving.config4_requests(ds, genes, n=n_req, seed=int(rate) + 7)
        sarr = serving.poisson_arrivals(n_req, rate=rate, seed=int(rate) + 8)
        _, slat, swall, ssizes = srv.run(sreqs, sarr)
        point = {"offered_rps": rate, "served_rps": round(n_req / swall, 1),
                 "p50_ms": round(float(np.percentile(slat, 50)) * 1e3, 3),
                 "p99_ms": round(float(np.percentile(slat, 99)) * 1e3, 3),
                 "mean_batch": round(float(np.mean(ssizes)), 1), "requests": n_req}
        sweep.append(point)
        log(f"capacity sweep: {point}")
        if point["p99_ms"] > args.p99_target_ms or point["served_rps"] < 0.95 * rate:
            break
        capacity = point
    mix = {k: sum(q["action"] == k for q in reqs) for k in ("mutations", "aggregate", "count")}
    # `value`: the largest offered rate served within the p99 bound (the fixed-rate run beside it)
    head = capacity or {"served_rps": qps, "p50_ms": float(np.percentile(latency, 50)) * 1e3,
                        "p99_ms": float(np.percentile(latency, 99)) * 1e3}
    line = {
        "metric": METRIC, "value": round(head["served_rps"], 3), "unit": "queries/s", "n_gpus": 1, "steps": 1, "warmup": 1,
        "ms_per_step": round(wall * 1e3, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u32", "data": "synthetic",
        "config": {"workload": (f"config 4: {ds.n_rows} sequences x {ds.genome_length} nt + {len(aa)} amino-acid columns, "
                                f"Poisson arrivals of the serving mix at rising rates, the largest one served with "
                                f"p99 <= {args.p99_target_ms:.0f} ms (and {len(reqs)} requests at {args.rate:.0f}/s)"),
                   "rows": ds.n_rows, "mix": mix, "path": "serving loop (continuous batching over lanes)"},
        "p50_ms": round(head["p50_ms"], 4),
        "p99_ms": round(head["p99_ms"], 4),
        "fixed_rate": {"offered_rps": args.rate, "served_rps": round(qps, 3),
                       "p50_ms": round(float(np.percentile(latency, 50)) * 1e3, 4),
                       "p99_ms": round(float(np.percentile(latency, 99)) * 1e3, 4)},
        "sparse_max_rows": tuned["sparse_max_rows"],
        "offered_rate": args.rate, "lanes": srv.lanes, "lane_MiB": round(srv.lane_bytes / 2**20, 1),
        "batches": len(sizes), "mean_batch": round(float(np.mean(sizes)), 2),
        "alone_equals_mix": {"sampled": len(sample), "identical": bool(identical)},
        "single_query_path_qps": round(plain_qps, 2), "upload_s": round(upload_s, 2),
        "capacity": {"p99_target_ms": args.p99_target_ms, "max_sustained_rps": capacity["offered_rps"] if capacity else None,
                     "at_capacity": capacity, "sweep": sweep},
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)
    et.close()
    eng.close()


def ingest(args):
    """--ingest: the index build of the config's main column (4 M sequences x 29,903 nt for config
    1), timed three ways: from the diff form with the host build; from the diff form with the
    local-reference adaptation and the sorted container entries on the GPU (the default); and from
    aligned sequence text through the column builder (diff on the GPU; the text is rendered on the
    host batch by batch, which is timed apart and subtracted).  All three layouts must be equal."""
    from silo_amd import datagen, table
    from silo_amd.engine import Engine

    t0 = time.time()
    ds = datagen.generate_config(args.config, n_rows=args.rows)
    spec = table.spec_from_dataset(ds)
    log(f"generated {ds.n_rows} rows x {ds.genome_length} nt in {time.time() - t0:.1f}s")
    main = spec.seq_columns[0]
    eng = Engine(0)

    def timed_build(col, device):
        eng.set_tuning("index_build_device", device)
        try:
            t = time.perf_counter()
            et = eng.table(table.TableSpec(chunk_sizes=spec.chunk_sizes, seq_columns=[col]))
            wall = time.perf_counter() - t
        finally:
            eng.set_tuning("index_build_device", 1)
        out = {"wall_s": round(wall, 3), "phases_ms": {k: round(v, 1) for k, v in et.build_profile("main").items()},
               "digest": f"{et.digest('main'):016x}"}
        et.close()
        return out

    host = timed_build(main, 0)
    log(f"host build: {host}")
    dev = timed_build(main, 1)
    log(f"device build: {dev}")
    render = [0.0]

    def batches():
        it = table.aligned_batches(main, 65536)
        while True:
            t = time.perf_counter()
            b = next(it, None)
            render[0] += time.perf_counter() - t
            if b is None:
                return
            yield b

    text_bytes = ds.n_rows * ds.genome_length
    aligned = timed_build(table.AlignedSeqColumnData("main", main.alphabet, main.reference, batches), 1)
    aligned["render_s"] = round(render[0], 3)
    aligned["builder_s"] = round(aligned["wall_s"] - render[0], 3)
    aligned["text_GB"] = round(text_bytes / 1e9, 2)
    aligned["text_GB_per_s"] = round(text_bytes / 1e9 / aligned["builder_s"], 2)
    log(f"aligned-text build: {aligned}")
    equal = host["digest"] == dev["digest"] == aligned["digest"]
    line = {
        "metric": "sequence index build time", "value": dev["wall_s"], "unit": "s", "n_gpus": 1, "steps": 1,
        "warmup": 0, "ms_per_step": round(dev["wall_s"] * 1e3, 1), "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "u32", "data": "synthetic",
        "config": {"workload": f"index build of {ds.n_rows} sequences x {ds.genome_length} nt (config {args.config})",
                   "rows": ds.n_rows},
        "host_build": host, "device_build": dev, "aligned_text_build": aligned, "layouts_equal": equal,
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)
    eng.close()
    if not equal:
        sys.exit(1)


def scan_pmc_bytes(kernel: str):
    """HBM bytes per launch of the batched dense scan from its FETCH_SIZE / WRITE_SIZE counter runs
    (profiles/batch_scan_pmc.json), when they cover this kernel; else None."""
    path = os.path.join(ROOT, "profiles", "batch_scan_pmc.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    return d.get("hbm_bytes_per_launch") if d.get("kernel", "").startswith(kernel) else None


def silo_load(args):
    """--silo: loading a saved database state (the reference's .silo archives) of the config's
    table.  The table (main column, date, country, lineage) is finalized on the host and written as
    a saved state first (untimed); the timed part reads the archives (schema + table), builds the
    upload form and uploads it (finalized-column build on the GPU).  The loaded table must answer
    the config's queries exactly as the same table uploaded in diff form."""
    import tempfile

    from silo_amd import datagen, silo_archive, table, workloads
    from silo_amd.engine import Engine
    from silo_amd.finalize import finalize_column

    t0 = time.time()
    ds = datagen.generate_config(args.config, n_rows=args.rows)
    spec = table.spec_from_dataset(ds)
    log(f"generated {ds.n_rows} rows x {ds.genome_length} nt in {time.time() - t0:.1f}s")
    t0 = time.perf_counter()
    fin = table.TableSpec(chunk_sizes=spec.chunk_sizes, int_columns=spec.int_columns, dict_columns=spec.dict_columns)
    fin.seq_columns = [finalize_column(c, spec.chunk_sizes) for c in spec.seq_columns]
    finalize_s = time.perf_counter() - t0
    root = tempfile.mkdtemp(prefix="silo_state_")
    t0 = time.perf_counter()
    path = silo_archive.save_database(root, {"default": fin}, int(time.time()))
    write_s = time.perf_counter() - t0
    del fin
    file_bytes = sum(os.path.getsize(os.path.join(path, f)) for f in os.listdir(path))
    log(f"finalized on the host in {finalize_s:.1f}s, wrote {file_bytes / 1e6:.1f} MB in {write_s:.1f}s")

    eng = Engine(0)
    t0 = time.perf_counter()
    db = silo_archive.load_database(path)
    read_s = time.perf_counter() - t0
    t1 = time.perf_counter()
    loaded_spec = db.table_spec()
    form_s = time.perf_counter() - t1
    t2 = time.perf_counter()
    loaded = eng.table(loaded_spec)
    upload_s = time.perf_counter() - t2
    total_s = time.perf_counter() - t0
    del db, loaded_spec
    log(f"read {read_s:.2f}s, upload form {form_s:.2f}s, upload {upload_s:.2f}s")
    t0 = time.perf_counter()
    direct = eng.table(spec)
    direct_s = time.perf_counter() - t0
    queries = workloads.config0_queries(ds) if args.config == 0 else workloads.config1_queries(ds)
    identical = True
    for q in queries + [{"filter": "(true)", "min_proportion": 0.05}]:
        f, p = q["filter"], q["min_proportion"]
        identical &= loaded.count(f) == direct.count(f)
        identical &= loaded.mutations_query(f, "main", p) == direct.mutations_query(f, "main", p)
    identical &= np.array_equal(loaded.local_reference("main"), direct.local_reference("main"))
    line = {
        "metric": "saved-state load time", "value": round(total_s, 3), "unit": "s", "n_gpus": 1, "steps": 1,
        "warmup": 0, "ms_per_step": round(total_s * 1e3, 1), "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "u32", "data": "synthetic",
        "config": {"workload": f"load of a saved state of {ds.n_rows} sequences x {ds.genome_length} nt "
                               f"(config {args.config}: main, date, country, lineage)", "rows": ds.n_rows},
        "archive_MB": round(file_bytes / 1e6, 1), "read_s": round(read_s, 3), "upload_form_s": round(form_s, 3),
        "upload_s": round(upload_s, 3), "read_MB_per_s": round(file_bytes / 1e6 / read_s, 1),
        "diff_form_upload_s": round(direct_s, 3), "host_finalize_s": round(finalize_s, 2),
        "write_s": round(write_s, 2), "queries_identical": bool(identical), "queries_checked": len(queries) + 1,
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)
    loaded.close()
    direct.close()
    eng.close()
    import shutil
    shutil.rmtree(root, ignore_errors=True)
    if not identical:
        sys.exit(1)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps (default: 200 for 64-query steps, 20 for the 1,024-tree steps of configs 2-3)")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", type=int, default=1)
    ap.add_argument("--rows", type=int, default=None)
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="one-thread CPU baseline leg (seconds)")
    ap.add_argument("--cpu-all-seconds", type=float, default=12.0, help="all-thread CPU baseline leg (seconds)")
    ap.add_argument("--cpu-threads", type=int, default=16,
                    help="threads of the all-thread CPU leg (at most the host's CPUs)")
    ap.add_argument("--steady-steps", type=int, default=1000,
                    help="pipelined steps of the steady-state figure reported beside the headline (0: skip)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-plain", action="store_true", help="skip the per-query path (the batched one still runs)")
    ap.add_argument("--no-batched", action="store_true", help="only the per-query path (for profiling its kernels)")
    ap.add_argument("--planner-trees", type=int, default=200000, help="config 3: training trees")
    ap.add_argument("--planner-eval", type=int, default=10000, help="config 3: held-out trees")
    ap.add_argument("--planner-epochs", type=int, default=30)
    ap.add_argument("--requests", type=int, default=4000, help="config 4: requests served")
    ap.add_argument("--rate", type=float, default=3000.0, help="config 4: Poisson arrival rate (requests/s)")
    ap.add_argument("--sweep-rates", default="2000,2500,3000,3500,4000,5000,6000,8000,10000,12000,16000,20000,24000,32000",
                    help="config 4: offered rates of the capacity sweep, ascending (empty: no sweep)")
    ap.add_argument("--sweep-seconds", type=float, default=4.0,
                    help="config 4: arrivals per sweep point, in seconds (the p99 of a heavy-tailed mix needs many)")
    ap.add_argument("--p99-target-ms", type=float, default=50.0, help="config 4: p99 latency bound of the sweep")
    ap.add_argument("--serving-graphs", type=int, default=1,
                    help="config 4: replay captured HIP graphs per lane job (0: launch the kernels directly)")
    ap.add_argument("--ingest", action="store_true", help="time the index build of the main column instead")
    ap.add_argument("--silo", action="store_true", help="time loading the config's table from a saved state instead")
    ap.add_argument("--knobs", default="",
                    help="key=value,... tuning knobs set instead of running the autotuner (e.g. for profiling)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB five times (scales PMC byte counters)")
    args = ap.parse_args()
    if args.steps is None:
        args.steps = 20 if args.config in (2, 3) else 200
    if args.gpus != 1:
        print("bench.py measures one GPU: this engine runs on a single MI355X (no multi-GPU path); "
              "use --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.ingest:
        ingest(args)
        return
    if args.silo:
        silo_load(args)
        return
    if args.config == 4:
        serve(args)
        return

    from silo_amd import autotune, datagen, table, workloads
    from silo_amd.engine import Engine

    t0 = time.time()
    ds = datagen.generate_config(args.config, n_rows=args.rows)
    spec = table.spec_from_dataset(ds)
    log(f"generated {ds.n_rows} rows x {ds.genome_length} nt in {time.time() - t0:.1f}s")

    eng = Engine(0)
    t0 = time.time()
    et = eng.table(spec)
    upload_s = time.time() - t0
    st = et.stats("main")
    log(f"index built and uploaded in {upload_s:.1f}s: {st}")

    if args.config == 0:
        queries = workloads.config0_queries(ds)
    elif args.config in (2, 3):
        t0 = time.time()

        def count(filters):
            out = []
            for i in range(0, len(filters), 2048):
                out += et.batch([{"filter": f, "action": "count"} for f in filters[i:i + 2048]])
            return [c if isinstance(c, int) else -1 for c in out]
        queries = workloads.config2_queries(ds, count)
        log(f"config-2 workload: {len(queries)} trees calibrated in {time.time() - t0:.1f}s")
    else:
        queries = workloads.config1_queries(ds)
    for q in queries:
        q.setdefault("action", "mutations")
        q.setdefault("column", "main")
    n_mut = sum(q["action"] == "mutations" for q in queries)
    log(f"{len(queries)} queries ({n_mut} Mutations)")

    # ---- plain path: one C-ABI call per query -----------------------------------------------------
    plain = None
    plain_results = None
    prof = None
    if not args.no_plain:
        def plain_pass(record=None):
            out = []
            for q in queries:
                t = time.perf_counter()
                out.append(run_plain_query(et, q))
                if record is not None:
                    record.append(time.perf_counter() - t)
            return out

        for _ in range(args.warmup):
            plain_pass()
        lat = []
        t_start = time.perf_counter()
        for _ in range(args.steps):
            plain_results = plain_pass(lat)
        elapsed = time.perf_counter() - t_start
        plain_results = as_rows(plain_results)
        plain = {"queries_per_s": round(args.steps * len(queries) / elapsed, 3),
                 "ms_per_step": round(elapsed / args.steps * 1e3, 3),
                 "p50_ms": round(percentile(lat, 50) * 1e3, 4), "p99_ms": round(percentile(lat, 99) * 1e3, 4)}
        # kernel durations: HIP events around every dense Mutations counting and filter launch, in a
        # pass of its own (the events make the host wait after each timed launch)
        eng.profile(True)
        plain_pass()
        prof = eng.profile_read()
        eng.profile(False)
        log(f"plain path: {plain}; kernel profile {prof}")

    # algorithmic bytes of the dense Mutations kernel (outside the timed regions)
    traffic = 0
    cards = []
    for q in queries:
        with et.bitmap(q["filter"]) as b:
            if q["action"] == "mutations":
                traffic += et.mutations_traffic(b, "main")
            cards.append(b.cardinality)

    if args.no_batched:
        n_dense = prof["mutations_launches"] if prof else 0
        print(json.dumps({"metric": METRIC, "value": plain["queries_per_s"] if plain else None, "unit": "queries/s",
                          "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "plain": plain,
                          "kernel_profile": prof, "dense_launches": n_dense,
                          "algorithmic_bytes_per_dense_launch": int(traffic / max(n_mut, 1))}), flush=True)
        et.close()
        eng.close()
        return

    # ---- batched path: autotune, then the whole step as one batch ---------------------------------
    t0 = time.time()
    tuned = None
    if args.knobs:
        # fixed knobs instead of the autotuner (profiling runs: the trace then holds the workload alone)
        for kv in args.knobs.split(","):
            k, v = kv.split("=")
            eng.set_tuning(k.strip(), int(v))
        log(f"fixed knobs: {args.knobs}")
    elif ds.n_rows >= 1 << 20:
        probes = autotune.probe_filters(ds)
        tuned = autotune.autotune(eng, et, probes, log=log)
        # the batched knobs on batches of the workload's own queries (the queries a server would see)
        mut = [q for q in queries if q["action"] == "mutations"]
        tuned.update(autotune.autotune_batch(eng, et, probes, log=log, cards=tuned["probe_cards"],
                                             workload=mut if len(mut) == len(queries) else None))
    log(f"autotune {time.time() - t0:.1f}s")
    # the plain path again with the autotuned single-query knobs (the sparse method below the
    # single-query switch): the batched path is also compared with this
    plain_tuned = None
    if tuned is not None and plain is not None:
        batch_switch = eng.get_tuning("sparse_max_rows")
        eng.set_tuning("sparse_max_rows", tuned["sparse_max_rows"])
        for q in queries:
            run_plain_query(et, q)
        t_pt = time.perf_counter()
        for _ in range(args.steps):
            pt_results = as_rows([run_plain_query(et, q) for q in queries])
        pt_elapsed = time.perf_counter() - t_pt
        eng.set_tuning("sparse_max_rows", batch_switch)
        plain_tuned = {"queries_per_s": round(args.steps * len(queries) / pt_elapsed, 3),
                       "sparse_max_rows": tuned["sparse_max_rows"],
                       "identical_to_plain": pt_results == plain_results}
        log(f"plain path with the tuned switch: {plain_tuned}")
    traffic_tuned = 0  # Mutations bytes by the method each query takes under the tuned switch
    for q in queries:
        if q["action"] == "mutations":
            with et.bitmap(q["filter"]) as b:
                traffic_tuned += et.mutations_traffic(b, "main")
    planner = None
    if args.config == 3:
        from silo_amd import planner_train

        t0 = time.time()
        train_trees = planner_train.trees(ds, args.planner_trees, seed=3)
        held_out = planner_train.trees(ds, args.planner_eval, seed=4)
        x, y = planner_train.samples(et, train_trees, log=log)
        t_samples = time.time() - t0
        mlp = eng.mlp(n_in=32, n_hidden=128, n_hidden_layers=3, seed=7)
        t0 = time.time()
        loss = None
        for ep in range(args.planner_epochs):  # one call per epoch (Adam state persists): a line each
            loss = mlp.train(x, y, epochs=1, batch=4096, lr=1e-3, seed=8 + ep)
            log(f"planner training: epoch {ep + 1}/{args.planner_epochs}, mse {loss:.5f}, {time.time() - t0:.1f}s")
        t_train = time.time() - t0
        ev = planner_train.evaluate(et, mlp, held_out)
        sp = planner_train.speedup(eng, et, queries, mlp)
        planner = {"training_trees": len(train_trees), "training_nodes": int(len(y)), "held_out": ev,
                   "train_mse": loss, "epochs": args.planner_epochs, "samples_s": round(t_samples, 1),
                   "train_s": round(t_train, 1), "workload": sp}
        log(f"planner: {planner}")
        eng.set_planner("learned", mlp)
    job = min(max(len(queries), 1), 256)
    if tuned and "batch_job_queries" in tuned and len(queries) <= tuned.get("probe_batch", 0):
        # a step the size of the autotuner's probe batch is cut into jobs as the probe's was
        job = max(1, len(queries) * tuned["batch_job_queries"] // tuned["probe_batch"])
    eng.set_tuning("batch_max_queries", job)
    for _ in range(max(args.warmup, 1)):
        et.batch(queries, arrays=True)
    # up to one batch per lane in flight: a step's host work (compiling and staging the next batch,
    # collecting an earlier one's results) overlaps the GPU work, and the kernels of batches in
    # flight share the GPU
    from collections import deque

    in_flight = max(1, eng.get_tuning("batch_lanes"))
    blat = []
    pend = deque()
    batch_results = None

    def collect():
        nonlocal batch_results
        t, b = pend.popleft()
        batch_results = b.results(arrays=True)
        b.free()
        blat.append(time.perf_counter() - t)

    t_start = time.perf_counter()
    for _ in range(args.steps):
        pend.append((time.perf_counter(), et.batch_submit(queries)))
        if len(pend) == in_flight:
            collect()
    while pend:
        collect()
    b_elapsed = time.perf_counter() - t_start
    batch_results = as_rows(batch_results)
    b_qps = args.steps * len(queries) / b_elapsed
    log(f"batched path: {b_qps:.1f} queries/s, {b_elapsed / args.steps * 1e3:.2f} ms per step")
    identical_paths = None
    if plain_results is not None:
        identical_paths = all(a == b for a, b in zip(plain_results, batch_results))
        log(f"batched results identical to plain: {identical_paths}")

    # steady state: the same pipelined loop over many more steps than the headline's K (its fill and
    # drain with several batches in flight weigh little here); reported beside `value`
    steady = None
    if args.steady_steps > 0:
        # at most about 10 s of steps (a 1,024-tree step takes 100 ms)
        per_step = b_elapsed / args.steps
        n_steady = max(16 * in_flight, min(args.steady_steps, int(10.0 / max(per_step, 1e-6))))
        pend = deque()
        t_st = time.perf_counter()
        for _ in range(n_steady):
            pend.append(et.batch_submit(queries))
            if len(pend) == in_flight:
                b = pend.popleft()
                b.results(arrays=True)
                b.free()
        while pend:
            b = pend.popleft()
            b.results(arrays=True)
            b.free()
        st_s = time.perf_counter() - t_st
        steady = {"queries_per_s": round(n_steady * len(queries) / st_s, 3), "steps": n_steady,
                  "ms_per_step": round(st_s / n_steady * 1e3, 3)}
        log(f"steady state: {steady}")

    # the batched path's dense counting kernel, timed with HIP events on its lane stream (jobs run
    # without graph replay while profiling), in a pass of its own outside the timed regions
    scan_roof = None
    if args.config == 1 and n_mut:
        eng.profile(True)
        for _ in range(3):
            et.batch(queries, arrays=True)
        sp = eng.profile_read()
        eng.profile(False)
        scan_roof = scan_roofline(eng, et, queries, sp, len(spec.chunk_sizes), log)

    # ---- CPU baseline ------------------------------------------------------------------------------
    cpu = None
    parity = None
    if not args.no_cpu_baseline:
        threads = max(1, min(args.cpu_threads, os.cpu_count() or 1))
        cpu, parity = cpu_legs(spec, queries, batch_results, args.cpu_seconds, args.cpu_all_seconds, threads)

    pmc = None
    pmc_path = os.path.join(ROOT, "profiles", "mutations_count_pmc.json")
    if os.path.exists(pmc_path):
        with open(pmc_path) as f:
            pmc = json.load(f).get("hbm_bytes_per_launch")
    roofline = None
    if prof is not None and prof["mutations_launches"] > 0:
        n_dense = prof["mutations_launches"]
        per_launch = traffic / max(n_mut, 1)
        achieved = per_launch * n_dense / (prof["mutations_ms"] * 1e-3) / 1e9
        # the device-side reference point: a plain streaming read of the same bytes, and of 4 GiB
        stream_same = eng.read_bandwidth(per_launch)
        stream_large = eng.read_bandwidth(4 << 30, reps=5)
        roofline = {"bound": "hbm", "kernel": "mutationsCountKernel", "achieved": round(achieved, 2),
                    "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": pmc,
                    "algorithmic_bytes_per_launch_mean": int(per_launch), "launches": n_dense,
                    "avg_launch_ms": round(prof["mutations_ms"] / n_dense, 5),
                    "stream_read_GBps_same_bytes": round(stream_same, 1), "stream_read_GBps_4GiB": round(stream_large, 1),
                    "frac_of_stream_read_same_bytes": round(achieved / stream_same, 4)}

    line = {
        "metric": METRIC,
        "value": round(b_qps, 3),
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(b_elapsed / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {
            "workload": (f"config {args.config}: {ds.n_rows} sequences x {ds.genome_length} nt, {len(queries)} queries "
                         f"per step ({n_mut} Mutations with minProportion 0.05)"),
            "rows": ds.n_rows,
            "genome_length": ds.genome_length,
            "queries_per_step": len(queries),
            "path": ("batched (one rdb_batch per step, up to one per lane in flight, autotuned knobs, HIP-graph replay"
                     + (", learned planner)" if args.config == 3 else ")")),
        },
        "p50_ms": round(percentile(blat, 50) * 1e3, 4),
        "p99_ms": round(percentile(blat, 99) * 1e3, 4),
        "latency_note": "batched path: a query's latency is its batch's submit-to-results time",
        "plain": plain,
        "batched_vs_plain": round(b_qps / plain["queries_per_s"], 3) if plain else None,
        "plain_tuned": plain_tuned,
        "batched_vs_plain_tuned": round(b_qps / plain_tuned["queries_per_s"], 3) if plain_tuned else None,
        "batched_identical_to_plain": identical_paths,
        "tuning": tuned,
        "planner": planner,
        "bitmap_GBps": round(traffic_tuned * args.steps / b_elapsed / 1e9, 2),
        "steady_state": steady,
        "roofline": roofline,
        "roofline_batched_scan": scan_roof,
        "cpu_baseline": cpu,
        "parity": parity,
        "upload_s": round(upload_s, 2),
        "filter_selectivity_range": [min(cards) / ds.n_rows, max(cards) / ds.n_rows],
    }
    print(json.dumps(line), flush=True)
    et.close()
    eng.close()


if __name__ == "__main__":
    main()
