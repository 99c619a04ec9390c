"""Uploading a sequence column in the state a finalized reference SequenceColumn holds
(rdb_table_add_sequence_column_finalized): the same table uploaded that way and in diff form gives
identical local references, index statistics, filter rows, counts, Mutations rows and groups, equal
to the oracle's; a local reference that adaptation from the reference would not choose (the
reference re-adapts after appends relative to the previous one) is kept; malformed containers are
rejected."""
import numpy as np
import pytest

from helpers import BITSET_FILTERS, ExprGen, bitset_table, small_spec

pytestmark = pytest.mark.gpu


def finalized_spec(spec, ot):
    from silo_amd import table

    out = table.TableSpec(chunk_sizes=spec.chunk_sizes, int_columns=spec.int_columns, dict_columns=spec.dict_columns,
                          bitmaps=spec.bitmaps)
    out.seq_columns = [ot.export_finalized(c.name) for c in spec.seq_columns]
    return out


def cases():
    from silo_amd import datagen, table

    ds, ts = small_spec(seed=13, n_rows=5000, L=500, chunk=2048)
    gen = ExprGen(ds, seed=5)
    yield "lineage", ts, ["(true)", "(has_mut main 17)"] + [gen.expr(3) for _ in range(40)], ["main"]
    spec, _, _ = bitset_table()
    yield "bitset", spec, BITSET_FILTERS, ["main"]
    ds = datagen.generate(datagen.GenSpec(n_rows=6000, genome_length=900, seed=77, lineage_model=True, n_nodes=40,
                                          n_mean=40, null_seq_frac=0.02), keep_handle=True)
    aa = datagen.translate(ds, datagen.gene_layout(ds.genome_length, [("g1", 120), ("g2", 150)], seed=5))
    gen = ExprGen(ds, seed=6)
    yield ("amino_acid", table.spec_from_dataset(ds, chunk=2048, aa_columns=aa),
           ["(true)", "(has_mut g1 30)", "(sym_eq g2 10 X)"] + [gen.expr(2) for _ in range(15)], ["main", "g1", "g2"])


@pytest.mark.parametrize("name,spec,filters,columns", list(cases()), ids=lambda x: x if isinstance(x, str) else "")
def test_finalized_upload_equals_diff_upload(engine, name, spec, filters, columns):
    from oracle.oracle import OracleTable

    ot = OracleTable(spec)
    a = engine.table(spec)
    b = engine.table(finalized_spec(spec, ot))
    try:
        for col in columns:
            assert np.array_equal(b.local_reference(col), a.local_reference(col))
            assert np.array_equal(b.local_reference(col), ot.local_reference(col))
            assert b.stats(col) == a.stats(col)
        for e in filters:
            rows = ot.filter(e)
            assert np.array_equal(b.filter(e), rows), e
            assert np.array_equal(a.filter(e), rows), e
        for e in filters[:12]:
            for col in columns:
                want = ot.mutations(e, col, 0.05)
                assert b.mutations_query(e, col, 0.05) == want, (e, col)
                assert np.array_equal(b.mutation_counts(e, col), ot.mutation_counts(e, col)), (e, col)
            dims = ["country", f"{columns[-1]}@3"] if name != "bitset" else ["country", "main@50"]
            assert b.aggregate(e, dims) == ot.aggregate(e, dims), e
        queries = [{"filter": e, "action": "mutations", "column": columns[0], "min_proportion": 0.02} for e in filters]
        assert b.batch(queries) == a.batch(queries)
    finally:
        a.close()
        b.close()
        ot.close()


def tie_column(local):
    """Position 0 with reference C over rows A, A, C, C (a tie: adaptation from C keeps C); position
    1 with reference G, all rows G.  The finalized state given has local reference `local` at 0."""
    from silo_amd import table

    A, C, G = 1, 2, 3
    n = 4
    if local == A:  # C's rows stored, A implicit
        keys, rows = [(0, 0, C)], [np.asarray([2, 3], np.uint16)]
    else:  # A's rows stored, C implicit
        keys, rows = [(0, 0, A)], [np.asarray([0, 1], np.uint16)]
    payload = np.concatenate([r.view(np.uint8) for r in rows])
    return table.FinalizedSeqColumnData(
        name="main", alphabet=0, reference=np.asarray([C, G], np.uint8), local_reference=np.asarray([local, G], np.uint8),
        key_position=np.asarray([k[0] for k in keys], np.uint32), key_chunk=np.asarray([k[1] for k in keys], np.uint16),
        key_symbol=np.asarray([k[2] for k in keys], np.uint8), typecode=np.asarray([2] * len(keys), np.uint8),
        cardinality=np.asarray([len(r) for r in rows], np.uint32), payload_offset=np.asarray([0], np.uint64),
        payload=payload, is_null=np.zeros(n, np.uint8), cov_start=np.zeros(n, np.uint32),
        cov_end=np.full(n, 2, np.uint32), missing_ptr=np.zeros(n + 1, np.uint64), missing_pos=np.zeros(0, np.uint32))


@pytest.mark.parametrize("local", [1, 2])
def test_given_local_reference_is_kept(engine, local):
    from silo_amd import table

    spec = table.TableSpec(chunk_sizes=np.asarray([4], np.uint32))
    spec.seq_columns.append(tie_column(local))
    et = engine.table(spec)
    try:
        assert et.local_reference("main").tolist() == [local, 3]
        assert et.filter("(sym_eq main 0 A)").tolist() == [0, 1]
        assert et.filter("(sym_eq main 0 C)").tolist() == [2, 3]
        assert et.filter("(has_mut main 0)").tolist() == [0, 1]
        assert et.mutations_query("(true)", "main", 0.0) == [(1, 2, 1, 2, 4, 0.5)]
        counts = et.mutation_counts("(true)", "main")
        assert counts[1, 0] == 2 and counts[2, 0] == 2 and counts[3, 1] == 4
    finally:
        et.close()


def test_malformed_finalized_columns_are_rejected(engine):
    from silo_amd import table
    from silo_amd.engine import EngineError

    def upload(col):
        spec = table.TableSpec(chunk_sizes=np.asarray([4], np.uint32))
        spec.seq_columns.append(col)
        with pytest.raises(EngineError) as err:
            engine.table(spec)
        assert err.value.status == "RDB_EINVAL"

    c = tie_column(1)
    c.key_symbol = np.asarray([1], np.uint8)  # a container of the local reference symbol
    upload(c)
    c = tie_column(1)
    c.key_chunk = np.asarray([1], np.uint16)  # a chunk outside the row layout
    upload(c)
    c = tie_column(1)
    c.cardinality = np.asarray([3], np.uint32)  # more values than the payload holds
    upload(c)
    c = tie_column(1)
    c.payload = np.asarray([7, 0, 3, 0], np.uint8)  # row 7 outside the chunk's 4 rows
    upload(c)
    c = tie_column(1)
    c.typecode = np.asarray([5], np.uint8)
    upload(c)
