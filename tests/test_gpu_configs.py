"""The benchmark workloads themselves on the GPU, at their full sizes.

* 100,000 sequences x 29,903 nt (the CPU-runnable plain case): every query of its workload
  compared with the oracle, on the single-query path by both Mutations methods and on every
  batched dense scan.
* 4,000,000 sequences x 29,903 nt from the 2,000-node lineage tree (61 full chunks and a
  2,304-row tail, 64 position blocks): the 64 Mutations queries of selectivity 1e-4..0.5 on the
  plain dense and sparse paths and every batched scan, all equal to each other, eight of them
  equal to the oracle; and the count identity sum_s count(s, p) + |F and not covered(p)| = |F|
  at sampled positions, whose second term comes from the filter kernel's coverage leaf
  (mutations_node.cpp:111-203 split a row between the symbols and the uncovered).
* 1,024 heterogeneous filter trees (depth 1-6 over And/Or/Not/NOf) on the same table, alternately
  Mutations and Aggregated by country: the batch equals the single-query path on all of them and
  the oracle on a sample; then the same batch under the learned and the independence plans
  returns the reference plan's results on every tree.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SCANS = [(0, 1024), (1, 1024), (2, 1024), (2, 512)]
KNOBS = ("sparse_max_rows", "batch_shared_scan", "batch_octet_threads", "batch_max_queries", "batch_lanes",
         "batch_row_capacity")


def log(msg):
    print(f"[configs] {msg}", flush=True)


@pytest.fixture
def tuning(engine):
    saved = {k: engine.get_tuning(k) for k in KNOBS}
    yield engine
    for k, v in saved.items():
        engine.set_tuning(k, v)
    engine.set_planner("reference")


def single(et, q):
    if q.get("action", "mutations") == "mutations":
        return et.mutations_query(q["filter"], q.get("column", "main"), q["min_proportion"])
    if q["action"] == "aggregate":
        with et.bitmap(q["filter"]) as b:
            return et.aggregate(b, q["dims"])
    return et.count(q["filter"])


def oracle(ot, q):
    if q.get("action", "mutations") == "mutations":
        return ot.mutations(q["filter"], q.get("column", "main"), q["min_proportion"])
    if q["action"] == "aggregate":
        return ot.aggregate(q["filter"], q["dims"])
    return ot.count(q["filter"])


def as_batch(queries):
    return [dict(q, action=q.get("action", "mutations"), column=q.get("column", "main")) for q in queries]


# ---- 100,000 rows ------------------------------------------------------------------------------------
@pytest.fixture(scope="module")
def config0(engine):
    from oracle.oracle import OracleTable
    from silo_amd import datagen, table

    ds = datagen.generate_config(0)
    spec = table.spec_from_dataset(ds)
    ot = OracleTable(spec)
    et = engine.table(spec)
    log(f"config 0 table: {ds.n_rows} rows")
    yield ds, spec, ot, et
    et.close()
    ot.close()


def test_config0_every_query(config0, tuning):
    from silo_amd import workloads

    ds, spec, ot, et = config0
    queries = as_batch(workloads.config0_queries(ds))
    want = [oracle(ot, q) for q in queries]
    counts = [ot.mutation_counts(q["filter"], "main") for q in queries]
    assert any(len(w) > 0 for w in want)
    for sparse in (-1, 1 << 40):
        tuning.set_tuning("sparse_max_rows", sparse)
        for q, w, c in zip(queries, want, counts):
            assert single(et, q) == w, (sparse, q)
            assert np.array_equal(et.mutation_counts(q["filter"], "main"), c), (sparse, q)
    tuning.set_tuning("sparse_max_rows", -1)
    for shared, threads in SCANS:
        tuning.set_tuning("batch_shared_scan", shared)
        tuning.set_tuning("batch_octet_threads", threads)
        assert et.batch(queries) == want, (shared, threads)
    tuning.set_tuning("sparse_max_rows", 2000)
    assert et.batch(queries) == want


# ---- 4,000,000 rows ----------------------------------------------------------------------------------
@pytest.fixture(scope="module")
def config1(engine):
    from oracle.oracle import OracleTable
    from silo_amd import datagen, table

    ds = datagen.generate_config(1)
    spec = table.spec_from_dataset(ds)
    et = engine.table(spec)
    log(f"config 1 table uploaded: {ds.n_rows} rows, {len(spec.chunk_sizes)} chunks")
    ot = OracleTable(spec)
    log("config 1 oracle built")
    state = {}
    yield ds, spec, ot, et, state
    et.close()
    ot.close()


def test_config1_geometry(config1):
    ds, spec, ot, et, state = config1
    assert len(spec.chunk_sizes) == 62 and int(spec.chunk_sizes[-1]) == 2304
    assert np.array_equal(et.local_reference("main"), ot.local_reference("main"))


def test_config1_plain_sparse_and_scans(config1, tuning):
    from silo_amd import workloads

    ds, spec, ot, et, state = config1
    queries = as_batch(workloads.config1_queries(ds))
    assert len(queries) == 64
    tuning.set_tuning("sparse_max_rows", -1)
    plain = [single(et, q) for q in queries]
    cards = [et.count(q["filter"]) for q in queries]
    sel = np.asarray(cards) / ds.n_rows
    assert sel.min() < 1e-3 and sel.max() > 0.3, sel
    sample = [0, 9, 18, 27, 36, 45, 54, 63]
    for i in sample:
        assert plain[i] == oracle(ot, queries[i]), (i, queries[i]["filter"])
        log(f"query {i} ({cards[i]} rows) equals the oracle")
    tuning.set_tuning("sparse_max_rows", 1 << 40)
    for i, q in enumerate(queries):
        assert single(et, q) == plain[i], ("sparse", i)
    tuning.set_tuning("batch_max_queries", 64)
    for sparse in (-1, 35000):
        tuning.set_tuning("sparse_max_rows", sparse)
        for shared, threads in SCANS:
            tuning.set_tuning("batch_shared_scan", shared)
            tuning.set_tuning("batch_octet_threads", threads)
            assert et.batch(queries) == plain, (sparse, shared, threads)
    log("batched scans equal the plain path")


def test_config1_count_identity(config1, tuning):
    """sum over symbols of count(s, p) plus the filter's rows not covered at p equals |F|, with the
    uncovered rows counted by the filter kernel (coverage leaf), not by the Mutations kernels."""
    from silo_amd import workloads

    ds, spec, ot, et, state = config1
    queries = workloads.config1_queries(ds)
    rng = np.random.default_rng(5)
    L = ds.genome_length
    for sparse in (-1, 1 << 40):
        tuning.set_tuning("sparse_max_rows", sparse)
        for i in (3, 30, 50, 62):
            f = queries[i]["filter"]
            counts = et.mutation_counts(f, "main").astype(np.int64)
            card = et.count(f)
            assert counts[15].sum() == 0  # the missing symbol is never counted
            pos = np.unique(np.r_[0, L - 1, rng.integers(0, L, 40), ds.nrun_start[:8].astype(np.int64) % L])
            unc = et.batch([{"filter": f"(and {f} (not (covered main {int(p)})))", "action": "count"} for p in pos])
            for p, u in zip(pos.tolist(), unc):
                assert int(counts[:, p].sum()) + u == card, (i, p)
            assert (counts.sum(axis=0) <= card).all()


def _config2_queries(config1):
    from silo_amd import workloads

    ds, spec, ot, et, state = config1
    if "queries" not in state:
        def count(filters):
            out = []
            for i in range(0, len(filters), 2048):
                out += et.batch([{"filter": f, "action": "count"} for f in filters[i:i + 2048]])
            return [c if isinstance(c, int) else -1 for c in out]
        state["queries"] = workloads.config2_queries(ds, count)
        log(f"config 2 workload: {len(state['queries'])} trees")
    return state["queries"]


def test_config2_trees(config1, tuning):
    ds, spec, ot, et, state = config1
    queries = _config2_queries(config1)
    assert len(queries) == 1024
    tuning.set_tuning("sparse_max_rows", 35000)
    tuning.set_tuning("batch_shared_scan", 1)
    tuning.set_tuning("batch_max_queries", 256)
    b = et.batch_submit(queries)
    try:
        got = b.results()
        cards = [b.count(i) for i in range(len(queries))]
    finally:
        b.free()
    state["reference_plan"] = got
    assert cards == [q["cardinality"] for q in queries]
    for i, q in enumerate(queries):
        assert got[i] == single(et, q), (i, q["filter"])
    log("1,024 trees: batch equals the single-query path")
    order = np.argsort([q["selectivity"] for q in queries], kind="stable")
    sample = [int(order[k]) for k in np.linspace(0, len(order) - 1, 12).astype(int)]
    for i in sample:
        assert got[i] == oracle(ot, queries[i]), (i, queries[i]["filter"])
    log(f"{len(sample)} sampled trees equal the oracle")


def test_config3_planned_results_unchanged(config1, tuning):
    from silo_amd import planner_train

    ds, spec, ot, et, state = config1
    queries = _config2_queries(config1)
    tuning.set_tuning("sparse_max_rows", 35000)
    tuning.set_tuning("batch_shared_scan", 1)
    tuning.set_tuning("batch_max_queries", 256)
    want = state.get("reference_plan") or et.batch(queries)
    x, y = planner_train.samples(et, planner_train.trees(ds, 1500, seed=3))
    mlp = tuning.mlp(n_in=32, n_hidden=128, n_hidden_layers=3, seed=7)
    try:
        mlp.train(x, y, epochs=4, batch=4096, lr=1e-3, seed=8)
        for mode in ("learned", "independence"):
            tuning.set_planner(mode, mlp if mode == "learned" else None)
            assert et.batch(queries) == want, mode
            log(f"{mode} plan: 1,024 results unchanged")
    finally:
        tuning.set_planner("reference")
        mlp.free()
