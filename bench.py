#!/usr/bin/env python3
"""Mutations-query benchmark on one MI355X.

A step is one pass of the configuration's query workload (config 1: 64 Mutations queries with
minProportion 0.05 over 4 M sequences x 29,903 nt; filters on date range, country or lineage with
selectivities 1e-4..0.5).  Inputs are generated, uploaded and resident in HBM before timing; each
query is filter + Mutations through the C ABI and returns its output rows to the host.  Prints one
JSON line.  The CPU baseline is the oracle (a single-threaded restatement of the reference's CPU
operators) timed on a bounded sample of the same queries, whose results are also compared with the
engine's bit for bit.
"""
import argparse
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "lapis-silo_amd"))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def percentile(xs, q):
    s = sorted(xs)
    k = min(len(s) - 1, max(0, int(round(q / 100.0 * (len(s) - 1)))))
    return s[k]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", type=int, default=1)
    ap.add_argument("--rows", type=int, default=None)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this engine runs on a single MI355X (no multi-GPU path); "
              "use --gpus 1", file=sys.stderr)
        sys.exit(2)

    from silo_amd import datagen, table, workloads
    from silo_amd.engine import Engine

    t0 = time.time()
    ds = datagen.generate_config(args.config, n_rows=args.rows)
    spec = table.spec_from_dataset(ds)
    queries = workloads.config0_queries(ds) if args.config == 0 else workloads.config1_queries(ds)
    log(f"generated {ds.n_rows} rows x {ds.genome_length} nt in {time.time() - t0:.1f}s; {len(queries)} queries")

    eng = Engine(0)
    t0 = time.time()
    et = eng.table(spec)
    upload_s = time.time() - t0
    st = et.stats("main")
    log(f"index built and uploaded in {upload_s:.1f}s: {st}")

    def run_pass(record=None):
        out = []
        for q in queries:
            t = time.perf_counter()
            rows = et.mutations_query(q["filter"], "main", q["min_proportion"])
            if record is not None:
                record.append(time.perf_counter() - t)
            out.append(rows)
        return out

    for _ in range(args.warmup):
        run_pass()
    eng.profile(True)
    lat = []
    t_start = time.perf_counter()
    results = None
    for _ in range(args.steps):
        results = run_pass(lat)
    elapsed = time.perf_counter() - t_start
    prof = eng.profile_read()
    eng.profile(False)
    n_q = args.steps * len(queries)
    qps = n_q / elapsed
    log(f"timed {n_q} queries in {elapsed:.3f}s: {qps:.1f} queries/s; kernel profile {prof}")

    # algorithmic bytes per query (outside the timed region)
    traffic = []
    cards = []
    for q in queries:
        with et.bitmap(q["filter"]) as b:
            traffic.append(et.mutations_traffic(b, "main"))
            cards.append(b.cardinality)
    bytes_per_pass = sum(traffic)
    mut_launch_bytes = bytes_per_pass * args.steps
    achieved = (mut_launch_bytes / (prof["mutations_ms"] * 1e-3)) / 1e9 if prof["mutations_ms"] > 0 else 0.0
    pmc = None
    pmc_path = os.path.join(ROOT, "profiles", "mutations_count_pmc.json")
    if os.path.exists(pmc_path):
        with open(pmc_path) as f:
            pmc = json.load(f).get("hbm_bytes_per_launch")

    cpu = None
    parity = None
    if not args.no_cpu_baseline:
        from oracle.oracle import OracleTable

        t0 = time.time()
        ot = OracleTable(spec)
        build_s = time.time() - t0
        log(f"oracle built in {build_s:.1f}s")
        done = 0
        identical = True
        t0 = time.perf_counter()
        while time.perf_counter() - t0 < args.cpu_seconds and done < len(queries):
            q = queries[(done * 7) % len(queries)]
            rows = ot.mutations(q["filter"], "main", q["min_proportion"])
            identical &= rows == results[(done * 7) % len(queries)]
            done += 1
        cpu_s = time.perf_counter() - t0
        cpu = {"value": round(done / cpu_s, 4), "unit": "queries/s", "cores": 1, "kind": "port",
               "sample": f"{done} of the {len(queries)} workload queries, single thread, oracle build "
                         f"{build_s:.1f}s excluded"}
        parity = {"queries_compared": done, "identical": bool(identical)}
        log(f"cpu baseline {cpu}; parity {parity}")
        ot.close()

    line = {
        "metric": "Mutations-query queries/s + p50/p99 latency @4M seqs, bit-exact; GPU vs CPU",
        "value": round(qps, 3),
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {
            "workload": (f"config {args.config}: {ds.n_rows} sequences x {ds.genome_length} nt, "
                         f"{len(queries)} Mutations(minProportion 0.05) queries per step"),
            "rows": ds.n_rows,
            "genome_length": ds.genome_length,
            "queries_per_step": len(queries),
            "path": "plain (per query: filter kernel + Mutations kernels, one host round trip)",
        },
        "p50_ms": round(percentile(lat, 50) * 1e3, 4),
        "p99_ms": round(percentile(lat, 99) * 1e3, 4),
        "bitmap_GBps": round(mut_launch_bytes / elapsed / 1e9, 2),
        "roofline": {
            "bound": "hbm",
            "kernel": "mutationsCountKernel",
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": pmc,
            "algorithmic_bytes_per_launch_mean": int(bytes_per_pass / len(queries)),
            "launches": prof["mutations_launches"],
            "avg_launch_ms": round(prof["mutations_ms"] / max(prof["mutations_launches"], 1), 5),
        },
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
