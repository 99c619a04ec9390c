"""Engine (HIP, through the C ABI) against the CPU oracle on the same seeded inputs: row bitmaps,
cardinalities, per-position mutation counts, Mutations rows and group counts must be identical."""
import numpy as np
import pytest

from helpers import ExprGen, random_bitmaps, small_spec

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def small(engine):
    from oracle.oracle import OracleTable

    ds, ts = small_spec(seed=11, n_rows=3000, L=400, chunk=1024)
    ts.bitmaps = random_bitmaps(ts, seed=5)
    ot = OracleTable(ts)
    et = engine.table(ts)
    yield ds, ts, ot, et
    et.close()
    ot.close()


def test_local_reference_matches(small):
    ds, ts, ot, et = small
    assert np.array_equal(et.local_reference("main"), ot.local_reference("main"))
    # the lineage model makes some positions adapt their local reference
    assert (ot.local_reference("main") != ds.reference).any()


def test_simple_filters(small):
    ds, ts, ot, et = small
    exprs = ["(true)", "(false)", "(sym_eq main 5 A)", "(has_mut main 17)", "(sym_eq main 33 N)",
             "(sym_eq main 33 .)", '(sym_in main 40 "AN")', "(covered main 12)", "(not (covered main 12))",
             f'(str_eq country "{ds.country_names[0]}")', "(str_eq country null)",
             f'(lineage lineage "{ds.lineage_names[0]}" 1)', "(date_between date 18262 18270)",
             "(bitmap 0)", "(not (bitmap 1))"]
    for e in exprs:
        assert np.array_equal(et.filter(e), ot.filter(e)), e
        assert et.count(e) == ot.count(e), e


def test_random_filters(small):
    ds, ts, ot, et = small
    gen = ExprGen(ds, seed=3, bitmaps=len(ts.bitmaps))
    for i in range(150):
        e = gen.expr(3)
        a = et.filter(e)
        b = ot.filter(e)
        assert np.array_equal(a, b), (i, e, len(a), len(b))


def test_mutations(small):
    ds, ts, ot, et = small
    gen = ExprGen(ds, seed=4)
    exprs = ["(true)", f'(str_eq country "{ds.country_names[0]}")', "(has_mut main 17)"] + [gen.expr(2) for _ in range(25)]
    for e in exprs:
        assert np.array_equal(et.mutation_counts(e, "main"), ot.mutation_counts(e, "main")), e
        for p in (0.0, 0.05, 0.5, 1.0):
            want = ot.mutations(e, "main", p)
            assert et.mutations(e, "main", p) == want, (e, p)
            assert et.mutations_query(e, "main", p) == want, (e, p)


def test_aggregate(small):
    ds, ts, ot, et = small
    gen = ExprGen(ds, seed=5)
    exprs = ["(true)", "(has_mut main 17)"] + [gen.expr(2) for _ in range(10)]
    for e in exprs:
        for dims in (["country"], ["lineage", "country"], ["main@33"], ["country", "main@5"], []):
            assert et.aggregate(e, dims) == ot.aggregate(e, dims), (e, dims)


@pytest.mark.parametrize("sparse_max", [0, 40, 1 << 40])
def test_mutations_sparse_switch(small, engine, sparse_max):
    """The row-walking method (filters of at most sparse_max rows) and the index-streaming one give
    the same counts and rows; 0 sends only empty filters to it, 2^40 sends every filter."""
    ds, ts, ot, et = small
    gen = ExprGen(ds, seed=9)
    exprs = ["(true)", "(false)", f'(str_eq country "{ds.country_names[0]}")', "(has_mut main 17)",
             "(nof 2 0 (has_mut main 17) (has_mut main 40) (sym_eq main 33 N))"] + [gen.expr(2) for _ in range(25)]
    engine.set_tuning("sparse_max_rows", sparse_max)
    try:
        assert engine.get_tuning("sparse_max_rows") == sparse_max
        for e in exprs:
            assert np.array_equal(et.mutation_counts(e, "main"), ot.mutation_counts(e, "main")), e
            for p in (0.0, 0.05, 1.0):
                want = ot.mutations(e, "main", p)
                assert et.mutations(e, "main", p) == want, (e, p)
                assert et.mutations_query(e, "main", p) == want, (e, p)
    finally:
        engine.set_tuning("sparse_max_rows", -1)


def test_tuning_keys(engine):
    from silo_amd.engine import EngineError

    with pytest.raises(EngineError):
        engine.set_tuning("no_such_knob", 1)
    with pytest.raises(EngineError):
        engine.set_tuning("mutation_groups", 33)
    assert engine.get_tuning("mutation_groups") == 0


def test_rejects_malformed_diffs(engine):
    """The engine takes the diff form aligned_sequence.cpp produces: mutations inside the covered
    range, outside missing runs and different from the reference."""
    from silo_amd import table
    from silo_amd.engine import EngineError

    ref = "ACGTACGTAC"
    for rows in (["ACGTACGTAC"], ["NNGTACGTAC"]):
        spec = table.TableSpec(chunk_sizes=np.asarray([1], np.uint32))
        spec.seq_columns.append(table.seq_column_from_strings("main", ref, rows, 0))
        engine.table(spec).close()
    bad = []
    for pos, sym in ((0, 3), (9, 3), (4, 1)):  # before the covered range, after it, equal to the reference
        c = table.seq_column_from_strings("main", ref, ["NNGTACGTNN"], 0)
        c.mut_ptr = np.asarray([0, 1], np.uint64)
        c.mut_pos = np.asarray([pos], np.uint32)
        c.mut_sym = np.asarray([sym], np.uint8)
        bad.append(c)
    c = table.seq_column_from_strings("main", ref, ["ACGTNNGTAC"], 0)  # inside the missing run [4, 6)
    c.mut_ptr = np.asarray([0, 1], np.uint64)
    c.mut_pos = np.asarray([5], np.uint32)
    c.mut_sym = np.asarray([1], np.uint8)
    bad.append(c)
    for c in bad:
        spec = table.TableSpec(chunk_sizes=np.asarray([1], np.uint32))
        spec.seq_columns.append(c)
        with pytest.raises(EngineError) as e:
            engine.table(spec)
        assert e.value.status == "RDB_EINVAL"


@pytest.mark.parametrize("sparse_max", [-1, 1 << 40])
def test_empty_covered_ranges(engine, sparse_max):
    """Rows with no covered position, whether given as [0, 0) or [L, L), count as missing
    everywhere on both Mutations methods."""
    from oracle.oracle import OracleTable
    from silo_amd import table

    ref = "ACGTACGTAC"
    rows = ["NNNNNNNNNN", "ACGTACGTAC", "ACGTTCGTAN", "NNGTACGTAC", "ACGAACGTAC"]
    spec = table.TableSpec(chunk_sizes=np.asarray([len(rows)], np.uint32))
    col = table.seq_column_from_strings("main", ref, rows, 0)
    col.cov_start = np.asarray(col.cov_start, np.uint32).copy()
    col.cov_end = np.asarray(col.cov_end, np.uint32).copy()
    col.cov_start[0] = col.cov_end[0] = len(ref)
    spec.seq_columns.append(col)
    ot = OracleTable(spec)
    et = engine.table(spec)
    engine.set_tuning("sparse_max_rows", sparse_max)
    try:
        for e in ("(true)", "(sym_eq main 3 A)", "(not (covered main 5))"):
            assert np.array_equal(et.mutation_counts(e, "main"), ot.mutation_counts(e, "main")), e
            assert et.mutations(e, "main", 0.0) == ot.mutations(e, "main", 0.0), e
    finally:
        engine.set_tuning("sparse_max_rows", -1)
        et.close()
        ot.close()


@pytest.mark.parametrize("coop", [0, 1])
def test_dense_kernel_shapes(engine, coop):
    """Every shape of the dense Mutations kernel (wavefronts sharing tiles or not, packed 16-bit or
    32-bit counters, loads in flight, chunk groups) gives the oracle's counts, on a table of 40
    chunks so that groups hold several chunks and the packed form its limit of 15."""
    from oracle.oracle import OracleTable
    from silo_amd import datagen, table

    spec = datagen.GenSpec(n_rows=40000, genome_length=300, seed=41, lineage_model=True, n_nodes=30, n_mean=20,
                           null_seq_frac=0.02, null_meta_frac=0.05, n_days=50)
    ds = datagen.generate(spec, keep_handle=True)
    ts = table.spec_from_dataset(ds, chunk=1024)
    ot = OracleTable(ts)
    et = engine.table(ts)
    keys = ("mutation_coop", "mutation_packed", "mutation_unroll", "mutation_groups", "sparse_max_rows")
    saved = {k: engine.get_tuning(k) for k in keys}
    exprs = ["(true)", "(has_mut main 17)", f'(str_eq country "{ds.country_names[1]}")',
             "(not (nof 2 0 (has_mut main 5) (has_mut main 40) (sym_eq main 77 N)))"]
    want = [ot.mutation_counts(e, "main") for e in exprs]
    try:
        engine.set_tuning("sparse_max_rows", -1)
        engine.set_tuning("mutation_coop", coop)
        for packed in (0, 1):
            for unroll in (8, 16):
                for groups in (0, 1, 3, 7, 32):
                    engine.set_tuning("mutation_packed", packed)
                    engine.set_tuning("mutation_unroll", unroll)
                    engine.set_tuning("mutation_groups", groups)
                    for e, w in zip(exprs, want):
                        assert np.array_equal(et.mutation_counts(e, "main"), w), (packed, unroll, groups, e)
    finally:
        for k, v in saved.items():
            engine.set_tuning(k, v)
        et.close()
        ot.close()


def test_int_columns_bit_sliced(engine):
    """Comparisons on int and date columns: bit-sliced copies (spans 0, 2,000 and 65,535, negative
    values, nulls) and the plain value scan (span 65,536) give the oracle's rows for thresholds
    inside, at the edges of and outside each column's range, with open bounds and negations."""
    from oracle.oracle import OracleTable
    from silo_amd import table

    rng = np.random.default_rng(17)
    n = 70000
    spec = table.TableSpec(chunk_sizes=np.asarray([40000, 30000], np.uint32))
    cols = {
        "a": (False, rng.integers(-1000, 1001, n)),
        "b": (False, np.concatenate([[7, 7 + 65535], rng.integers(7, 7 + 65536, n - 2)])),
        "c": (False, np.concatenate([[-5, -5 + 65536], rng.integers(-5, -5 + 65537, n - 2)])),
        "d": (False, np.full(n, 42)),
        "e": (True, rng.integers(18000, 19001, n)),
    }
    for name, (is_date, vals) in cols.items():
        nulls = (rng.random(n) < 0.05).astype(np.uint8)
        nulls[:2] = 0  # keep the range ends
        spec.int_columns.append(table.IntColumnData(name, is_date, vals.astype(np.int32), nulls))
    ot = OracleTable(spec)
    et = engine.table(spec)
    try:
        exprs = []
        for name, (is_date, vals) in cols.items():
            lo, hi = int(vals.min()), int(vals.max())
            mid = int(vals[rng.integers(0, n)])
            points = [lo - 1, lo, lo + 1, mid, hi - 1, hi, hi + 1, -2**31, 2**31 - 1]
            op = "date_between" if is_date else "int_between"
            for v in points:
                if not is_date:
                    exprs.append(f"(int_eq {name} {v})")
                exprs += [f"({op} {name} {v} _)", f"({op} {name} _ {v})", f"(not ({op} {name} {v} {mid}))"]
            exprs += [f"({op} {name} {lo} {hi})", f"({op} {name} {mid} {hi + 5})", f"({op} {name} _ _)"]
        for e in exprs:
            assert np.array_equal(et.filter(e), ot.filter(e)), e
    finally:
        et.close()
        ot.close()
