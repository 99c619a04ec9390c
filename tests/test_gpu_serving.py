"""The serving loop: requests of the config-4 mix (nucleotide and amino-acid Mutations, Aggregated
by lineage x country, n-of-m counts) arrive while earlier batches are on the GPU; each result
collected from the mix equals the result of the same query run alone (single-query path) and the
oracle's, bit for bit."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GENES = [("g1", 240), ("g2", 300), ("g3", 150)]


@pytest.fixture(scope="module")
def served(engine):
    from oracle.oracle import OracleTable
    from silo_amd import datagen, table

    ds = datagen.generate(datagen.GenSpec(n_rows=20000, genome_length=3000, seed=77, lineage_model=True, n_nodes=60,
                                          n_mean=40, null_seq_frac=0.01, null_meta_frac=0.02), keep_handle=True)
    aa = datagen.translate(ds, datagen.gene_layout(ds.genome_length, GENES, seed=5))
    ts = table.spec_from_dataset(ds, chunk=8192, aa_columns=aa)
    ot = OracleTable(ts)
    et = engine.table(ts)
    yield ds, ts, ot, et
    et.close()
    ot.close()


def alone(et, q):
    if q["action"] == "mutations":
        return et.mutations_query(q["filter"], q["column"], q["min_proportion"])
    if q["action"] == "aggregate":
        with et.bitmap(q["filter"]) as b:
            return et.aggregate(b, q["dims"])
    return et.count(q["filter"])


def oracle(ot, q):
    if q["action"] == "mutations":
        return ot.mutations(q["filter"], q["column"], q["min_proportion"])
    if q["action"] == "aggregate":
        return ot.aggregate(q["filter"], q["dims"])
    return ot.count(q["filter"])


@pytest.mark.parametrize("realtime", [True, False])
def test_results_alone_equal_results_in_the_mix(served, engine, realtime):
    from silo_amd import serving

    ds, ts, ot, et = served
    saved = {k: engine.get_tuning(k) for k in ("batch_lanes", "batch_max_queries", "batch_row_capacity",
                                               "sparse_max_rows")}
    try:
        engine.set_tuning("sparse_max_rows", 300)
        reqs = serving.config4_requests(ds, [g for g, _ in GENES], n=240, seed=11, lo=1e-4)
        arrivals = serving.poisson_arrivals(len(reqs), rate=20000.0, seed=12)
        srv = serving.Server(engine, et, max_batch=24, row_capacity=4096, max_lanes=3)
        assert srv.lanes >= 1
        results, latency, wall, sizes = srv.run(reqs, arrivals, realtime=realtime)
        assert sum(sizes) == len(reqs) and len(sizes) > 1
        assert np.all(latency >= 0)
        kinds = {q["action"] for q in reqs}
        assert kinds == {"mutations", "aggregate", "count"}
        assert any(q.get("column", "main") != "main" for q in reqs if q["action"] == "mutations")
        for q, r in zip(reqs, results):
            assert r == alone(et, q), q
            assert r == oracle(ot, q), q
    finally:
        for k, v in saved.items():
            engine.set_tuning(k, v)
