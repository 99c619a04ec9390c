"""Bitset containers on the GPU: a table of full 65,536-row chunks whose vertical index and
dictionary bitmaps hold bitsets (helpers.bitset_table) gives the oracle's filter rows, counts,
Mutations rows and groups through every kernel that reads a bitset: the filter kernel's sequence and
store leaves, the single-query dense Mutations kernel in each shape, the sparse method, and every
batched dense scan (one, four and eight queries per pass).  The counting loop pinned here is the
reference's container AND-cardinality (mutations_node.cpp:154-189)."""
import numpy as np
import pytest

from helpers import BITSET_FILTERS, bitset_table

pytestmark = pytest.mark.gpu

SCANS = [(0, 1024), (1, 1024), (2, 1024), (2, 512)]
MUT_FILTERS = ["(true)", "(sym_eq main 10 T)", '(str_eq country "c1")', "(not (sym_eq main 31 -))",
               "(has_mut main 20)", "(sym_eq main 50 C)", "(sym_eq main 61 N)",
               "(nof 2 0 (has_mut main 10) (has_mut main 20) (has_mut main 50) (sym_eq main 62 C))"]
KNOBS = ("sparse_max_rows", "mutation_coop", "mutation_packed", "mutation_unroll", "mutation_groups",
         "batch_shared_scan", "batch_octet_threads", "batch_max_queries")


@pytest.fixture(scope="module")
def data(engine):
    from oracle.oracle import OracleTable

    spec, M, null = bitset_table()
    print(f"bitset table: {spec.n_rows} rows", flush=True)
    ot = OracleTable(spec)
    et = engine.table(spec)
    saved = {k: engine.get_tuning(k) for k in KNOBS}
    yield spec, ot, et
    for k, v in saved.items():
        engine.set_tuning(k, v)
    et.close()
    ot.close()


def test_bitset_filter_leaves(data):
    spec, ot, et = data
    assert np.array_equal(et.local_reference("main"), ot.local_reference("main"))
    for e in BITSET_FILTERS:
        assert np.array_equal(et.filter(e), ot.filter(e)), e
        assert et.count(e) == ot.count(e), e


@pytest.mark.parametrize("coop", [0, 1])
def test_bitset_dense_kernel_shapes(data, engine, coop):
    spec, ot, et = data
    want = [ot.mutation_counts(e, "main") for e in MUT_FILTERS]
    engine.set_tuning("sparse_max_rows", -1)
    engine.set_tuning("mutation_coop", coop)
    for packed in (0, 1):
        for unroll in (8, 16):
            for groups in (0, 1, 2):
                engine.set_tuning("mutation_packed", packed)
                engine.set_tuning("mutation_unroll", unroll)
                engine.set_tuning("mutation_groups", groups)
                for e, w in zip(MUT_FILTERS, want):
                    assert np.array_equal(et.mutation_counts(e, "main"), w), (packed, unroll, groups, e)


@pytest.mark.parametrize("sparse_max", [-1, 1 << 40])
def test_bitset_mutations_rows(data, engine, sparse_max):
    spec, ot, et = data
    engine.set_tuning("sparse_max_rows", sparse_max)
    for e in MUT_FILTERS:
        for p in (0.0, 0.05):
            want = ot.mutations(e, "main", p)
            assert et.mutations_query(e, "main", p) == want, (e, p)


@pytest.mark.parametrize("shared,threads", SCANS)
def test_bitset_batched_scans(data, engine, shared, threads):
    spec, ot, et = data
    engine.set_tuning("sparse_max_rows", -1)
    engine.set_tuning("batch_shared_scan", shared)
    engine.set_tuning("batch_octet_threads", threads)
    engine.set_tuning("batch_max_queries", 16)
    queries = [{"filter": e, "action": "mutations", "column": "main", "min_proportion": p}
               for e in MUT_FILTERS for p in (0.0, 0.05)]
    got = et.batch(queries)
    for q, g in zip(queries, got):
        assert g == ot.mutations(q["filter"], "main", q["min_proportion"]), q


def test_bitset_aggregate(data):
    spec, ot, et = data
    for e in ["(true)", "(has_mut main 20)", '(str_eq country "c2")']:
        for dims in (["country"], ["main@50"], ["country", "main@10"]):
            assert et.aggregate(e, dims) == ot.aggregate(e, dims), (e, dims)
