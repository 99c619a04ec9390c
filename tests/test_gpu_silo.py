"""Loading the reference's saved database state onto the GPU.

The saved state of testBaseData/unitTestDummyDataset (database.test.cpp:71-96), read by
silo_amd.silo_archive and uploaded through the finalized-column path, answers filters, Count,
Mutations and Aggregated queries exactly as the same rows built from input.ndjson in diff form do,
and as the brute-force oracle does on the aligned strings.  Lineage filters resolve aliases as the
reference's lineage index does (lineage_index.cpp:17-45)."""
import numpy as np
import pytest

import silo_support as ss

pytestmark = pytest.mark.gpu

FILTERS = [
    "(true)", "(false)",
    '(lineage pango_lineage "B.1.1.7" 0)', '(lineage pango_lineage "B.1.1" 0)', '(lineage pango_lineage "B.1.1" 1)',
    '(lineage pango_lineage "B" 1)', '(lineage pango_lineage "XBB" 1)', '(lineage pango_lineage "A" 1)',
    '(str_eq country "Switzerland")', '(str_eq division "Bern")', '(str_eq division "Basel-Land")',
    "(date_between date 18704 18730)", "(date_between unsorted_date _ 18650)", "(date_between unsorted_date 18600 _)",
    "(sym_eq main 1 A)", "(sym_eq main 3 N)", "(sym_eq main 7 .)", "(has_mut main 5)", "(covered main 7)",
    "(sym_eq testSecondSequence 3 N)", "(has_mut M 0)", "(sym_eq M 2 X)", '(sym_in E 0 "MX")',
    '(and (lineage pango_lineage "B.1.1" 1) (not (sym_eq main 3 N)))',
    '(or (str_eq division "Bern") (date_between date 18742 _))',
    '(nof 2 0 (has_mut main 1) (sym_eq M 0 X) (lineage pango_lineage "XBB" 0))',
]
COLUMNS = ["main", "testSecondSequence", "E", "M"]


def brute_table(spec):
    """The brute-force oracle over the dataset's aligned strings and the spec's metadata."""
    from oracle import brute

    rs = ss.rows()
    seqs = {name: (alphabet, ref, [(r.get(name) or {}).get("sequence") for r in rs])
            for name, (alphabet, ref) in ss.references().items()}
    ints = {c.name: (c.values, c.is_null) for c in spec.int_columns}
    dicts = {c.name: (c.ids, c.is_null, c.names, c.lineage_parent, c.lineage_alias) for c in spec.dict_columns}
    return brute.BruteTable(spec.chunk_sizes, seq_columns=seqs, int_columns=ints, dict_columns=dicts)


@pytest.fixture(scope="module")
def tables(engine, tmp_path_factory):
    from silo_amd import silo_archive as sa

    db = sa.load_database(ss.saved_state(tmp_path_factory.mktemp("state")))
    loaded_spec = db.table_spec()
    text = db.tables["default"].schema.column("pango_lineage").lineage_tree.file
    built_spec = ss.ndjson_spec(text)
    loaded = engine.table(loaded_spec)
    built = engine.table(built_spec)
    yield loaded, built, brute_table(loaded_spec), brute_table(built_spec), loaded_spec, built_spec
    loaded.close()
    built.close()


def test_loaded_table_filters_like_the_ndjson_rows(tables):
    loaded, built, bl, bb = tables[:4]
    for e in FILTERS:
        want = bb.filter(e)
        assert np.array_equal(bl.filter(e), want), e
        assert np.array_equal(loaded.filter(e), want), e
        assert np.array_equal(built.filter(e), want), e
        assert loaded.count(e) == len(want), e
    assert loaded.filter('(lineage pango_lineage "B" 1)').tolist() == [0, 1, 2]
    assert loaded.filter('(str_eq division "Bern")').tolist() == [1, 3]


def named(groups, spec, dims):
    """Groups with dictionary ids replaced by their strings (the two tables number values apart)."""
    names = {c.name: c.names for c in spec.dict_columns}
    out = []
    for key, count in groups:
        out.append((tuple(v if v is None or d not in names else names[d][v] for d, v in zip(dims, key)), count))
    return out


def test_loaded_table_mutations_and_groups(tables):
    loaded, built, bl, bb, loaded_spec, built_spec = tables
    for col in COLUMNS:
        assert np.array_equal(loaded.local_reference(col), built.local_reference(col)), col
        for e in FILTERS[:10]:
            assert np.array_equal(loaded.mutation_counts(e, col), built.mutation_counts(e, col)), (e, col)
            for p in (0.0, 0.05, 0.5):
                want = bb.mutations(e, col, p)
                assert loaded.mutations_query(e, col, p) == want, (e, col, p)
                assert built.mutations_query(e, col, p) == want, (e, col, p)
    # key2's AAGNAAGN over ACGTACGT: C>A at positions 2 and 6 (1-based), 1 of 5 covered rows each
    assert loaded.mutations_query("(true)", "main", 0.0) == [(2, 2, 1, 1, 5, 0.2), (6, 2, 1, 1, 5, 0.2)]
    for dims in (["country"], ["division", "pango_lineage"], ["pango_lineage", "main@3"]):
        for e in ("(true)", '(lineage pango_lineage "B.1.1" 1)', "(date_between date 18705 _)"):
            want = named(bb.aggregate(e, dims), built_spec, dims)
            assert named(bl.aggregate(e, dims), loaded_spec, dims) == want, (e, dims)
            assert named(loaded.aggregate(e, dims), loaded_spec, dims) == want, (e, dims)
            assert named(built.aggregate(e, dims), built_spec, dims) == want, (e, dims)
    queries = [{"filter": e, "action": "mutations", "column": c, "min_proportion": 0.05} for e in FILTERS for c in COLUMNS]
    assert loaded.batch(queries) == built.batch(queries)


def test_lineage_aliases_resolve_like_the_lineage_index(engine):
    """A.1 has the alias X.1: rows stored as X.1 are A.1's rows in lineage filters (and A's
    sublineage rows); a filter on X.1 is a filter on A.1; string equality keeps the stored value."""
    from oracle import brute
    from silo_amd import table

    names = ["A", "A.1", "B", "X.1", "A.1.1"]
    parent = np.array([-1, 0, -1, -1, 1], np.int32)
    alias = np.array([-1, -1, -1, 1, -1], np.int32)
    ids = np.array([0, 1, 3, 3, 2, 1, 4, 0], np.uint32)
    nulls = np.array([0, 0, 0, 0, 0, 0, 0, 1], np.uint8)
    spec = table.TableSpec(chunk_sizes=np.array([len(ids)], np.uint32))
    spec.dict_columns.append(table.DictColumnData("lineage", ids, nulls, names, parent, alias))
    bt = brute.BruteTable(spec.chunk_sizes, dict_columns={"lineage": (ids, nulls, names, parent, alias)})
    et = engine.table(spec)
    try:
        want = {
            '(lineage lineage "A.1" 0)': [1, 2, 3, 5],
            '(lineage lineage "X.1" 0)': [1, 2, 3, 5],
            '(lineage lineage "X.1" 1)': [1, 2, 3, 5, 6],
            '(lineage lineage "A" 1)': [0, 1, 2, 3, 5, 6],
            '(lineage lineage "A" 0)': [0],
            '(lineage lineage "B" 1)': [4],
            '(str_eq lineage "X.1")': [2, 3],
            '(str_eq lineage "A.1")': [1, 5],
            "(lineage lineage null 0)": [7],
        }
        for e, rows in want.items():
            assert bt.filter(e).tolist() == rows, e
            assert et.filter(e).tolist() == rows, e
    finally:
        et.close()


def test_saved_generated_table_answers_like_its_diff_form(engine, tmp_path):
    """A generated table written as a saved state (host finalize + writer) and loaded answers every
    filter, Mutations and Aggregated query as the same table uploaded in diff form, and as the
    oracle."""
    from helpers import ExprGen, small_spec
    from oracle.oracle import OracleTable
    from silo_amd import silo_archive as sa, table
    from silo_amd.finalize import finalize_column

    ds, ts = small_spec(seed=23, n_rows=7000, L=700, chunk=2048)
    spec = table.TableSpec(chunk_sizes=ts.chunk_sizes, int_columns=ts.int_columns, dict_columns=ts.dict_columns)
    spec.seq_columns = [finalize_column(c, ts.chunk_sizes) for c in ts.seq_columns]
    loaded_spec = sa.load_database(sa.save_database(str(tmp_path), {"default": spec}, 7)).table_spec()
    gen = ExprGen(ds, seed=9)
    filters = ["(true)", "(has_mut main 17)"] + [gen.expr(3) for _ in range(30)]
    ot = OracleTable(ts)
    a = engine.table(ts)
    b = engine.table(loaded_spec)
    try:
        assert np.array_equal(b.local_reference("main"), ot.local_reference("main"))
        for e in filters:
            rows = ot.filter(e)
            assert np.array_equal(b.filter(e), rows), e
        for e in filters[:10]:
            assert b.mutations_query(e, "main", 0.05) == ot.mutations(e, "main", 0.05), e
            assert a.mutations_query(e, "main", 0.05) == ot.mutations(e, "main", 0.05), e
            assert b.aggregate(e, ["country", "main@3"]) == a.aggregate(e, ["country", "main@3"]), e
        queries = [{"filter": e, "action": "mutations", "column": "main", "min_proportion": 0.02} for e in filters]
        assert b.batch(queries) == a.batch(queries)
    finally:
        a.close()
        b.close()
        ot.close()


def test_saved_amino_acid_and_bitset_tables(engine, tmp_path):
    """Amino-acid columns and 65,536-row chunks with bitset and run containers survive a saved state:
    the loaded tables answer Mutations on every column like the oracle."""
    from helpers import BITSET_FILTERS, bitset_table
    from oracle.oracle import OracleTable
    from silo_amd import datagen, silo_archive as sa, table
    from silo_amd.finalize import finalize_column

    ds = datagen.generate(datagen.GenSpec(n_rows=6000, genome_length=900, seed=78, lineage_model=True, n_nodes=40,
                                          n_mean=40, null_seq_frac=0.02), keep_handle=True)
    aa = datagen.translate(ds, datagen.gene_layout(ds.genome_length, [("g1", 120), ("g2", 150)], seed=5))
    cases = [(table.spec_from_dataset(ds, chunk=2048, aa_columns=aa), ["(true)", "(has_mut g1 30)", "(sym_eq g2 10 X)",
                                                                       "(has_mut main 200)"], ["main", "g1", "g2"]),
             (bitset_table()[0], BITSET_FILTERS, ["main"])]
    for k, (ts, filters, columns) in enumerate(cases):
        spec = table.TableSpec(chunk_sizes=ts.chunk_sizes, int_columns=ts.int_columns, dict_columns=ts.dict_columns)
        spec.seq_columns = [finalize_column(c, ts.chunk_sizes) for c in ts.seq_columns]
        loaded = sa.load_database(sa.save_database(str(tmp_path / str(k)), {"default": spec}, 11)).table_spec()
        ot = OracleTable(ts)
        et = engine.table(loaded)
        try:
            for col in columns:
                assert np.array_equal(et.local_reference(col), ot.local_reference(col)), col
            for e in filters:
                assert np.array_equal(et.filter(e), ot.filter(e)), e
                for col in columns:
                    assert et.mutations_query(e, col, 0.02) == ot.mutations(e, col, 0.02), (e, col)
        finally:
            et.close()
            ot.close()
