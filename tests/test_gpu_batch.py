"""The batched path (rdb_batch_*) against the oracle: every query of a batch, whatever its action,
lane job, dense or sparse method and graph replay, returns what the oracle returns for it alone."""
import numpy as np
import pytest

from helpers import ExprGen, random_bitmaps, small_spec

pytestmark = pytest.mark.gpu

KNOBS = ("sparse_max_rows", "batch_max_queries", "batch_lanes", "batch_graphs", "batch_row_capacity", "batch_shared_scan",
         "batch_octet_threads", "batch_octet_groups", "batch_fail_job")
# (batch_shared_scan, batch_octet_threads): every dense counting kernel of a batch
SCANS = [(0, 1024), (1, 1024), (2, 1024), (2, 512)]


@pytest.fixture(scope="module")
def small(engine):
    from oracle.oracle import OracleTable

    ds, ts = small_spec(seed=21, n_rows=5000, L=500, chunk=2048)
    ts.bitmaps = random_bitmaps(ts, seed=6)
    ot = OracleTable(ts)
    et = engine.table(ts)
    yield ds, ts, ot, et
    et.close()
    ot.close()


@pytest.fixture
def tuning(engine):
    saved = {k: engine.get_tuning(k) for k in KNOBS}
    yield engine
    for k, v in saved.items():
        engine.set_tuning(k, v)


def make_queries(ds, seed, n, bitmaps=0):
    gen = ExprGen(ds, seed=seed, bitmaps=bitmaps)
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        e = gen.expr(int(rng.integers(1, 4)))
        kind = i % 4
        if kind == 0:
            out.append({"filter": e, "action": "count"})
        elif kind in (1, 2):
            out.append({"filter": e, "action": "mutations", "column": "main",
                        "min_proportion": float(rng.choice([0.0, 0.02, 0.05, 0.5, 1.0]))})
        else:
            dims = [["country"], ["lineage", "country"], []][int(rng.integers(0, 3))]
            out.append({"filter": e, "action": "aggregate", "dims": dims})
    return out


def expected(ot, q):
    if q["action"] == "count":
        return ot.count(q["filter"])
    if q["action"] == "mutations":
        return ot.mutations(q["filter"], q["column"], q["min_proportion"])
    return ot.aggregate(q["filter"], q["dims"])


def check(et, ot, queries):
    got = et.batch(queries)
    assert len(got) == len(queries)
    for q, g in zip(queries, got):
        assert g == expected(ot, q), q


@pytest.mark.parametrize("shared,threads", SCANS)
@pytest.mark.parametrize("graphs", [0, 1])
@pytest.mark.parametrize("sparse_max", [-1, 300, 1 << 40])
def test_batch_matches_oracle(small, tuning, graphs, sparse_max, shared, threads):
    ds, ts, ot, et = small
    tuning.set_tuning("batch_graphs", graphs)
    tuning.set_tuning("batch_shared_scan", shared)
    tuning.set_tuning("batch_octet_threads", threads)
    tuning.set_tuning("sparse_max_rows", sparse_max)
    tuning.set_tuning("batch_max_queries", 16)  # several lane jobs per batch
    check(et, ot, make_queries(ds, seed=100 + graphs, n=60, bitmaps=len(ts.bitmaps)))


def test_batch_replays_stay_exact(small, tuning):
    """The same captured graph replayed for several batches, with single-query work between them."""
    ds, ts, ot, et = small
    tuning.set_tuning("batch_graphs", 1)
    tuning.set_tuning("sparse_max_rows", 500)
    tuning.set_tuning("batch_max_queries", 32)
    queries = make_queries(ds, seed=7, n=32)
    want = [expected(ot, q) for q in queries]
    for rep in range(4):
        assert et.batch(queries) == want, rep
        assert et.count("(has_mut main 17)") == ot.count("(has_mut main 17)")
    other = make_queries(ds, seed=8, n=20)
    assert et.batch(other) == [expected(ot, q) for q in other]
    assert et.batch(queries) == want


def test_batch_overlapping_submissions(small, tuning):
    """Batches submitted before earlier ones finished (more jobs than lanes) keep their results."""
    ds, ts, ot, et = small
    tuning.set_tuning("batch_lanes", 2)
    tuning.set_tuning("batch_max_queries", 8)
    tuning.set_tuning("sparse_max_rows", 200)
    qsets = [make_queries(ds, seed=200 + k, n=20) for k in range(4)]
    handles = [et.batch_submit(qs) for qs in qsets]
    try:
        # collected newest first: older batches' jobs were retired by later submissions
        for qs, h in zip(reversed(qsets), reversed(handles)):
            assert h.results() == [expected(ot, q) for q in qs]
    finally:
        for h in handles:
            h.free()


def test_batch_row_overflow_reruns(small, tuning):
    ds, ts, ot, et = small
    tuning.set_tuning("batch_row_capacity", 16)
    queries = [{"filter": "(true)", "action": "mutations", "column": "main", "min_proportion": 0.0},
               {"filter": "(has_mut main 17)", "action": "mutations", "column": "main", "min_proportion": 0.0}]
    check(et, ot, queries)


def test_batch_query_errors_are_per_query(small):
    from silo_amd.engine import EngineError

    ds, ts, ot, et = small
    queries = [{"filter": "(has_mut main 17)", "action": "count"},
               {"filter": "(sym_eq nosuch 3 A)", "action": "count"},
               {"filter": "(true)", "action": "mutations", "column": "nosuch", "min_proportion": 0.05},
               {"filter": "(true", "action": "count"},
               {"filter": "(true)", "action": "mutations", "column": "main", "min_proportion": 1.5},
               {"filter": "(true)", "action": "aggregate", "dims": ["main@3"]},
               {"filter": "(str_eq country null)", "action": "aggregate", "dims": ["lineage"]}]
    for arrays in (False, True):
        got = et.batch(queries, arrays=arrays)
        assert got[0] == ot.count("(has_mut main 17)")
        for g in got[1:6]:
            assert isinstance(g, EngineError) and g.status == "RDB_EQUERY", g
        assert got[6] == ot.aggregate("(str_eq country null)", ["lineage"])


def test_shared_scan_many_chunks(engine, tuning):
    """More than 2 x 15 chunks: the four-query scan's halves and the eight-query scan's quarters each
    gather several chunk runs in 16-bit counters and flush them more than once; results equal the
    per-query kernel's and the oracle's."""
    from oracle.oracle import OracleTable

    ds, ts = small_spec(seed=31, n_rows=40000, L=300, chunk=1024)
    ot = OracleTable(ts)
    et = engine.table(ts)
    try:
        tuning.set_tuning("sparse_max_rows", -1)
        tuning.set_tuning("batch_max_queries", 16)
        queries = [{"filter": f, "action": "mutations", "column": "main", "min_proportion": mp}
                   for f, mp in [("(true)", 0.0), ("(true)", 0.05), ("(not (has_mut main 17))", 0.01),
                                 ("(str_eq country \"%s\")" % ds.country_names[0], 0.0),
                                 ("(nof 1 0 (has_mut main 5) (has_mut main 40) (has_mut main 77))", 0.02),
                                 ("(date_between date %d %d)" % (ds.spec.date_base, ds.spec.date_base + 20), 0.0)]]
        want = [ot.mutations(q["filter"], "main", q["min_proportion"]) for q in queries]
        for shared, threads in SCANS:
            tuning.set_tuning("batch_shared_scan", shared)
            tuning.set_tuning("batch_octet_threads", threads)
            assert et.batch(queries) == want, (shared, threads)
        # the persistent eight-query scan with fewer work-groups than CUs (each takes more items)
        tuning.set_tuning("batch_shared_scan", 2)
        tuning.set_tuning("batch_octet_threads", 512)
        for groups in (1, 7, 64):
            tuning.set_tuning("batch_octet_groups", groups)
            assert et.batch(queries) == want, groups
    finally:
        et.close()


def test_full_chunk_run(engine, tuning):
    """A non-reference symbol that holds every row of a full 65,536-row chunk (one run container of
    65,536 rows; the symbol is still a minority of the column): its count under a filter of all rows
    is 65,536 in every dense kernel, the single-query ones and each batched scan."""
    from oracle.oracle import OracleTable
    from silo_amd import table

    ref = "ACGTACGT"
    rows = ["ACGTTCGT"] * 65536 + ["ACGTACGT"] * 65536 + ["ACGTACGA"] * 65536
    spec = table.TableSpec(chunk_sizes=np.asarray([65536, 65536, 65536], np.uint32))
    spec.seq_columns.append(table.seq_column_from_strings("main", ref, rows, 0))
    ot = OracleTable(spec)
    et = engine.table(spec)
    try:
        tuning.set_tuning("sparse_max_rows", -1)
        queries = [{"filter": f, "action": "mutations", "column": "main", "min_proportion": 0.0}
                   for f in ("(true)", "(sym_eq main 4 T)", "(not (sym_eq main 7 A))")]
        want = [ot.mutations(q["filter"], "main", 0.0) for q in queries]
        assert any(r[3] == 65536 for r in want[0]), want[0]
        for q, w in zip(queries, want):
            assert et.mutations(q["filter"], "main", 0.0) == w, q
        for shared, threads in SCANS:
            tuning.set_tuning("batch_shared_scan", shared)
            tuning.set_tuning("batch_octet_threads", threads)
            assert et.batch(queries) == want, (shared, threads)
    finally:
        et.close()
        ot.close()
        ot.close()


def test_results_as_columns(small, tuning):
    """Mutations results delivered as columns hold the same rows as the tuple form, on both paths."""
    from silo_amd.engine import MutationArrays

    ds, ts, ot, et = small
    queries = make_queries(ds, seed=300, n=24)
    rows = et.batch(queries)
    cols = et.batch(queries, arrays=True)
    for q, r, c in zip(queries, rows, cols):
        if q["action"] == "mutations":
            assert isinstance(c, MutationArrays) and len(c) == 6
            assert c.rows() == r, q
            single = et.mutations_query(q["filter"], q["column"], q["min_proportion"], arrays=True)
            assert single.rows() == r, q
        else:
            assert c == r, q


def test_failed_submit_leaves_lanes_usable(small, tuning):
    """A lane job that fails after others of its batch were enqueued fails that submission only:
    a batch already in flight keeps its results, and every lane serves later batches."""
    from silo_amd.engine import EngineError

    ds, ts, ot, et = small
    tuning.set_tuning("batch_lanes", 3)
    tuning.set_tuning("batch_max_queries", 8)
    tuning.set_tuning("sparse_max_rows", 200)
    first = make_queries(ds, seed=400, n=16)
    pending = et.batch_submit(first)  # two jobs in flight
    try:
        tuning.set_tuning("batch_fail_job", 2)  # the second job of the next submission fails
        with pytest.raises(EngineError) as err:
            et.batch_submit(make_queries(ds, seed=401, n=40))
        assert err.value.status == "RDB_ENOMEM"
        assert tuning.get_tuning("batch_fail_job") == 0
        assert pending.results() == [expected(ot, q) for q in first]
    finally:
        pending.free()
        tuning.set_tuning("batch_fail_job", 0)
    # all three lanes in use at once, results exact
    qsets = [make_queries(ds, seed=410 + k, n=8) for k in range(3)]
    handles = [et.batch_submit(qs) for qs in qsets]
    try:
        for qs, h in zip(qsets, handles):
            assert h.results() == [expected(ot, q) for q in qs]
    finally:
        for h in handles:
            h.free()


@pytest.mark.parametrize("n_chunks,effective", [(4100, 1), (8200, 0)])
def test_shared_scan_falls_back_on_many_chunks(engine, tuning, n_chunks, effective):
    """The eight-query scan holds at most 1,024 chunks per part and the four-query scan 4,096 per
    half: on larger tables a job runs the next smaller scan, the lane holds (and its footprint
    counts) only that scan's buffers, and results equal the oracle's."""
    from oracle.oracle import OracleTable
    from silo_amd import datagen, table

    ds = datagen.generate(datagen.GenSpec(n_rows=n_chunks * 3, genome_length=60, seed=57, lineage_model=True,
                                          n_nodes=20, n_mean=5, null_seq_frac=0.02, n_days=30), keep_handle=True)
    ts = table.spec_from_dataset(ds, chunk=3)
    assert len(ts.chunk_sizes) == n_chunks
    ot = OracleTable(ts)
    et = engine.table(ts)
    try:
        tuning.set_tuning("sparse_max_rows", -1)
        tuning.set_tuning("batch_max_queries", 8)
        tuning.set_tuning("batch_shared_scan", effective)
        reference_bytes = et.lane_bytes(8, 4096)
        tuning.set_tuning("batch_shared_scan", 2)
        assert et.lane_bytes(8, 4096) == reference_bytes
        queries = [{"filter": f, "action": "mutations", "column": "main", "min_proportion": mp}
                   for f, mp in [("(true)", 0.0), ("(has_mut main 7)", 0.05),
                                 (f'(str_eq country "{ds.country_names[0]}")', 0.0),
                                 ("(not (sym_eq main 30 N))", 0.02)]]
        want = [ot.mutations(q["filter"], "main", q["min_proportion"]) for q in queries]
        assert et.batch(queries) == want
    finally:
        et.close()
        ot.close()
