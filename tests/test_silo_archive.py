"""Reading the reference's saved database state (CPU only).

The fixture is the state database.test.cpp:71-96 saves from testBaseData/unitTestDummyDataset.
The loader must read the whole of each archive, give the counts that test checks (5 sequences,
a non-empty vertical index, 9 bytes of horizontal bitmaps), and hold exactly what the dataset's
input.ndjson says: each sequence column equal to the oracle's finalized column of the same rows,
the int, date, bool and dictionary values row by row, and the lineage tree of the lineage
definition text the schema carries.  Damaged or foreign archives are refused."""
import gzip
import os
import struct

import numpy as np
import pytest

import silo_support as ss


@pytest.fixture(scope="module")
def db(tmp_path_factory):
    from silo_amd import silo_archive as sa

    root = os.path.dirname(ss.saved_state(tmp_path_factory.mktemp("state")))
    path = sa.most_recent_data_directory(root)
    assert path is not None and os.path.basename(path) == ss.TIMESTAMP
    return sa.load_database(path)


def test_database_info_matches_the_reference_test(db):
    # database.test.cpp:97-112: sequence_count 5, vertical_bitmaps_size > 0, horizontal_bitmaps_size 9
    from silo_amd import silo_archive as sa

    t = db.tables["default"]
    assert t.sequence_count == 5 and t.chunk_sizes.tolist() == [5]
    cols = [c for c in t.columns.values() if isinstance(c, sa.SequenceColumnState)]
    assert sorted(c.name for c in cols) == ["E", "M", "main", "testSecondSequence"]
    assert all(c.info[0] == 5 and c.sequence_count == 5 for c in cols)
    assert sum(c.info[1] for c in cols) > 0
    assert sum(c.info[2] for c in cols) == 9
    assert db.timestamp == int(ss.TIMESTAMP) and db.serialization_version == sa.SERIALIZATION_VERSION


def test_schema_lists_the_dataset_columns(db):
    from silo_amd import silo_archive as sa

    s = db.tables["default"].schema
    assert s.primary_key == "primaryKey"
    types = {c.name: c.type for c in s.columns}
    assert types["main"] == sa.NUC_SEQUENCE and types["E"] == sa.AA_SEQUENCE
    assert types["date"] == sa.DATE32 and types["age"] == sa.INT32 and types["qc_value"] == sa.FLOAT
    assert types["pango_lineage"] == sa.DICTIONARY_ENCODED and types["test_boolean_column"] == sa.BOOL
    assert types["unaligned_main"] == sa.ZSTD_STRING and types["primaryKey"] == sa.STRING
    refs = ss.references()
    for name, (alphabet, ref) in refs.items():
        from silo_amd import table

        want = [table.symbol_of(alphabet, ch) for ch in ref]
        assert s.column(name).reference.tolist() == want


def container_rows(f, k):
    from silo_amd import silo_archive as sa

    o = int(f.payload_offset[k])
    return sa.container_values(int(f.typecode[k]), int(f.cardinality[k]), bytes(f.payload[o:o + 8192 + 4]))


def index_sets(f) -> dict:
    return {(int(f.key_position[k]), int(f.key_chunk[k]), int(f.key_symbol[k])): container_rows(f, k).tolist()
            for k in range(len(f.key_position))}


def test_sequence_columns_equal_the_oracle_finalize_of_the_ndjson_rows(db):
    from oracle.oracle import OracleTable

    spec = db.table_spec()
    lineage = db.tables["default"].schema.column("pango_lineage").lineage_tree
    ot = OracleTable(ss.ndjson_spec(lineage.file, aliases=False))
    try:
        for got in spec.seq_columns:
            want = ot.export_finalized(got.name)
            assert got.local_reference.tolist() == want.local_reference.tolist(), got.name
            assert index_sets(got) == index_sets(want), got.name
            assert all(int(got.cardinality[k]) == len(container_rows(got, k)) for k in range(len(got.key_position)))
            for a in ("is_null", "cov_start", "cov_end"):
                assert getattr(got, a).tolist() == getattr(want, a).tolist(), (got.name, a)
            for x, y in zip(got.missing_positions(), want.missing_positions()):
                assert x.tolist() == y.tolist(), got.name
    finally:
        ot.close()
    main = next(c for c in spec.seq_columns if c.name == "main")
    # row key2 = AAGNAAGN over ACGTACGT: C>A at 1 and 5, N at 3 inside the covered range, N at 7 ends it
    assert index_sets(main) == {(1, 0, 1): [1], (5, 0, 1): [1]}
    assert main.cov_end.tolist() == [8, 7, 8, 8, 8] and main.missing_positions()[1].tolist() == [3]


def test_metadata_columns_hold_the_ndjson_values(db):
    t = db.tables["default"]
    spec = db.table_spec()
    rs = ss.rows()
    ints = {c.name: c for c in spec.int_columns}
    for name in ("date", "unsorted_date"):
        vals = [r.get(name) for r in rs]
        assert ints[name].is_date
        assert ints[name].is_null.tolist() == [int(v is None) for v in vals]
        assert [x for x, v in zip(ints[name].values.tolist(), vals) if v is not None] == \
               [ss.days(v) for v in vals if v is not None]
    assert ints["age"].values.tolist() == [r["age"] for r in rs] and not ints["age"].is_date
    dicts = {c.name: c for c in spec.dict_columns}
    for name in ("country", "division", "region", "pango_lineage"):
        c = dicts[name]
        assert [c.names[i] for i in c.ids.tolist()] == [r[name] for r in rs], name
    b = t.columns["test_boolean_column"]
    flags = [r["test_boolean_column"] for r in rs]
    assert b["true_rows"].tolist() == [i for i, v in enumerate(flags) if v is True]
    assert b["false_rows"].tolist() == [i for i, v in enumerate(flags) if v is False]
    assert b["null_rows"].tolist() == [i for i, v in enumerate(flags) if v is None]
    q = t.columns["qc_value"]
    assert np.concatenate(q["chunks"]).tolist() == [r["qc_value"] for r in rs]
    assert not any(d["null_rows"].size for n, d in t.columns.items() if n in ("age", "qc_value", "country"))


def test_lineage_tree_is_the_definition_it_was_built_from(db):
    from silo_amd import silo_archive as sa

    col = db.tables["default"].schema.column("pango_lineage")
    tree = col.lineage_tree
    assert col.dictionary == tree.names
    names, parent, alias = ss.lineage_definition(tree.file)
    got_parent, got_alias = sa.lineage_arrays(tree, len(col.dictionary))
    assert sorted(names) == sorted(col.dictionary)
    index = {n: i for i, n in enumerate(names)}
    perm = np.array([index[n] for n in col.dictionary])
    # the same tree up to the order of ids: parents and alias targets by name
    by_name = lambda arr, ns: [None if x < 0 else ns[x] for x in arr.tolist()]
    assert by_name(got_parent, col.dictionary) == [by_name(parent, names)[i] for i in perm]
    assert by_name(got_alias, col.dictionary) == [by_name(alias, names)[i] for i in perm]
    # the saved lineage index: rows under each of B.1.1.7's ancestors, including itself
    li = db.tables["default"].columns["pango_lineage"]["lineage_index"]
    b117 = col.dictionary.index("B.1.1.7")
    chain = []
    v = b117
    while v >= 0:
        chain.append(v)
        v = int(got_parent[v])
    for mode, index_including in li["including_sublineages"].items():
        assert sorted(k for k, rows in index_including.items() if rows.values().tolist() == [0, 1, 2]) == sorted(chain)


def encode_portable(values: np.ndarray, runs: bool) -> bytes:
    """CRoaring's portable format, written from its published specification (RoaringFormatSpec):
    cookie, container keys and cardinalities, offsets, then array / bitset / run containers."""
    values = np.unique(np.asarray(values, np.uint32))
    keys = sorted(set((values >> 16).tolist()))
    conts = []
    for k in keys:
        low = (values[(values >> 16) == k] & 0xFFFF).astype(np.uint32)
        breaks = np.nonzero(np.diff(low) != 1)[0]
        starts = np.concatenate([[low[0]], low[breaks + 1]])
        ends = np.concatenate([low[breaks], [low[-1]]])
        if runs and len(starts) * 4 + 2 < min(2 * len(low), 8192):
            body = struct.pack("<H", len(starts)) + b"".join(struct.pack("<HH", s, e - s) for s, e in zip(starts, ends))
            conts.append((k, len(low), True, body))
        elif len(low) > 4096:
            bits = np.zeros(65536, np.uint8)
            bits[low] = 1
            conts.append((k, len(low), False, np.packbits(bits, bitorder="little").tobytes()))
        else:
            conts.append((k, len(low), False, low.astype("<u2").tobytes()))
    n = len(conts)
    has_run = any(c[2] for c in conts)
    if has_run:
        flags = bytearray((n + 7) // 8)
        for i, c in enumerate(conts):
            if c[2]:
                flags[i // 8] |= 1 << (i % 8)
        head = struct.pack("<I", 12347 | ((n - 1) << 16)) + bytes(flags)
    else:
        head = struct.pack("<II", 12346, n)
    head += b"".join(struct.pack("<HH", k, card - 1) for k, card, _, _ in conts)
    if not has_run or n >= 4:
        off = len(head) + 4 * n
        for c in conts:
            head += struct.pack("<I", off)
            off += len(c[3])
    return head + b"".join(c[3] for c in conts)


@pytest.mark.parametrize("runs", [False, True])
def test_portable_roaring_decode_round_trip(runs):
    from silo_amd import silo_archive as sa

    rng = np.random.default_rng(3)
    for trial in range(20):
        parts = [rng.integers(0, 1 << 16, size=int(rng.integers(1, 50)), dtype=np.uint32) + (7 << 16),
                 np.arange(3 << 16, (3 << 16) + int(rng.integers(1, 65536)), dtype=np.uint32),
                 rng.choice(1 << 16, size=int(rng.integers(4097, 30000)), replace=False).astype(np.uint32) + (9 << 16)]
        vals = np.concatenate(parts[: 1 + trial % 3])
        got = sa.roaring_values(encode_portable(vals, runs))
        assert got.tolist() == np.unique(vals).tolist()
    assert sa.roaring_values(struct.pack("<II", 12346, 0)).size == 0


def _runs_of(values: np.ndarray) -> list:
    """Maximal runs [a, b) of sorted unique values, by brute force."""
    out = []
    for v in values.tolist():
        if out and out[-1][1] == v:
            out[-1][1] = v + 1
        else:
            out.append([v, v + 1])
    return [tuple(r) for r in out]


def test_native_bitmap_runs_match_the_values():
    """rdb_archive_bitmap_runs (the reader's decode of horizontal bitmaps) gives the maximal runs of
    the values each bitmap holds, for array, bitset and run containers, several containers per
    bitmap and runs that cross a container boundary; malformed bitmaps raise SiloFormatError."""
    from silo_amd import silo_archive as sa

    rng = np.random.default_rng(11)
    bitmaps, want = [], []
    for trial in range(60):
        parts = [rng.integers(0, 1 << 16, size=int(rng.integers(1, 50)), dtype=np.uint32) + (7 << 16),
                 np.arange((3 << 16) - int(rng.integers(0, 9)), (3 << 16) + int(rng.integers(1, 65536)), dtype=np.uint32),
                 rng.choice(1 << 16, size=int(rng.integers(4097, 30000)), replace=False).astype(np.uint32) + (9 << 16),
                 np.zeros(0, np.uint32)]
        vals = np.unique(np.concatenate(parts[: 1 + trial % 4]))
        bitmaps.append(encode_portable(vals, trial % 2 == 1))
        want.append(_runs_of(vals))
    bitmaps.append(struct.pack("<II", 12346, 0))
    want.append([])
    data = np.frombuffer(b"".join(bitmaps), np.uint8)
    starts = np.cumsum([0] + [len(b) for b in bitmaps[:-1]])
    ptr, rs, re_ = sa.decode_bitmap_runs(data, starts, np.array([len(b) for b in bitmaps]))
    for i, w in enumerate(want):
        got = list(zip(rs[ptr[i]:ptr[i + 1]].tolist(), re_[ptr[i]:ptr[i + 1]].tolist()))
        assert got == w, i
    good = encode_portable(np.array([1, 2, 3, 70000], np.uint32), False)
    for bad in (good[:-1], b"\x00" * 8, good[:4], struct.pack("<II", 12346, 70000)):
        with pytest.raises(sa.SiloFormatError):
            sa.decode_bitmap_runs(np.frombuffer(bad, np.uint8), np.array([0]), np.array([len(bad)]))


def test_damaged_or_foreign_archives_are_refused(tmp_path):
    from silo_amd import silo_archive as sa

    d = ss.saved_state(tmp_path)
    with open(os.path.join(d, "default.silo"), "rb") as f:
        table_bytes = f.read()
    with open(os.path.join(d, "database_schema.silo"), "rb") as f:
        schema = sa.read_schema(f.read())["default"]
    for bad, what in ((table_bytes[:-3], "truncated"), (table_bytes + b"\0", "trailing"),
                      (b"\x05" + table_bytes[1:], "not a Boost"), (table_bytes[:30] + b"\x09\x00" + table_bytes[32:], "library version")):
        with pytest.raises(sa.SiloFormatError, match=what):
            sa.read_table(bad, schema)
    with open(os.path.join(d, "data_version.silo"), "w") as f:
        f.write("timestamp: 1785915539\nserializationVersion: 1\n")
    with pytest.raises(sa.SiloFormatError, match="serialization version"):
        sa.load_database(d)
    assert sa.most_recent_data_directory(os.path.dirname(d)) is None


def test_writer_reproduces_the_saved_sequence_columns_byte_for_byte(db, tmp_path):
    """A sequence column written back from what was read gives the archive's own bytes (columns with
    insertions aside: the engine does not hold insertions)."""
    from silo_amd import silo_archive as sa

    d = ss.saved_state(tmp_path)
    with open(os.path.join(d, "default.silo"), "rb") as f:
        raw = f.read()
    t = db.tables["default"]
    spec = db.table_spec()
    compared = 0
    for c in spec.seq_columns:
        st = t.columns[c.name]
        if st.insertion_count:
            continue
        w = sa.ArchiveWriter()
        base = len(w.buf)
        w.classes = set(st.classes_before)
        sa._write_sequence_column(w, c, spec.chunk_sizes)
        assert bytes(w.buf[base:]) == raw[st.span[0]:st.span[1]], c.name
        compared += 1
    assert compared == 3


def same(a, b) -> bool:
    for f in a.__dataclass_fields__:
        x, y = getattr(a, f), getattr(b, f)
        if isinstance(x, np.ndarray) or isinstance(y, np.ndarray):
            if (x is None) != (y is None) or (x is not None and not np.array_equal(np.asarray(x), np.asarray(y))):
                return False
        elif x != y:
            return False
    return True


def test_saved_state_round_trip(db, tmp_path):
    """The fixture's tables written and read again are the same upload form; the lineage index the
    writer computes from parents and aliases is the one the reference saved."""
    from silo_amd import silo_archive as sa

    spec = db.table_spec()
    d = sa.save_database(str(tmp_path), {"default": spec}, 1234)
    assert sa.most_recent_data_directory(str(tmp_path)) == d
    back = sa.load_database(d)
    spec2 = back.table_spec()
    for a, b in zip(spec.seq_columns + spec.int_columns + spec.dict_columns,
                    spec2.seq_columns + spec2.int_columns + spec2.dict_columns):
        assert same(a, b), a.name
    li = db.tables["default"].columns["pango_lineage"]["lineage_index"]
    li2 = back.tables["default"].columns["pango_lineage"]["lineage_index"]
    flat = lambda m: {k: v.values().tolist() for k, v in m.items()}
    assert flat(li["excluding_sublineages"]) == flat(li2["excluding_sublineages"])
    for mode in range(3):
        assert flat(li["including_sublineages"][mode]) == flat(li2["including_sublineages"][mode])


def generated_tables():
    from helpers import bitset_table, small_spec
    from silo_amd import datagen, table

    yield "lineage", small_spec(seed=21, n_rows=5000, L=600, chunk=1500)[1]
    yield "bitset", bitset_table()[0]
    ds = datagen.generate(datagen.GenSpec(n_rows=6000, genome_length=900, seed=78, lineage_model=True, n_nodes=40,
                                          n_mean=40, null_seq_frac=0.02), keep_handle=True)
    aa = datagen.translate(ds, datagen.gene_layout(ds.genome_length, [("g1", 120), ("g2", 150)], seed=5))
    yield "amino_acid", table.spec_from_dataset(ds, chunk=2048, aa_columns=aa)


@pytest.mark.parametrize("name,ts", list(generated_tables()), ids=lambda x: x if isinstance(x, str) else "")
def test_generated_table_round_trip(tmp_path, name, ts):
    """Generated tables (partial chunks, nulls, adapted local references, missing runs; 65,536-row
    chunks with bitset and run containers; amino-acid columns) finalized on the host, saved and
    loaded give back the same finalized columns and metadata."""
    from silo_amd import silo_archive as sa, table
    from silo_amd.finalize import finalize_column

    spec = table.TableSpec(chunk_sizes=ts.chunk_sizes, int_columns=ts.int_columns, dict_columns=ts.dict_columns)
    spec.seq_columns = [finalize_column(c, ts.chunk_sizes) for c in ts.seq_columns]
    if name == "lineage":
        assert (spec.seq_columns[0].local_reference != spec.seq_columns[0].reference).any()
    if name == "bitset":
        assert {1, 2, 3} <= set(np.unique(spec.seq_columns[0].typecode).tolist())
    d = sa.save_database(str(tmp_path), {"default": spec}, 99)
    back = sa.load_database(d).table_spec()
    assert back.chunk_sizes.tolist() == spec.chunk_sizes.tolist()
    for a, b in zip(sorted(spec.seq_columns, key=lambda c: c.name), sorted(back.seq_columns, key=lambda c: c.name)):
        assert same(a, b)
    for a, b in zip(sorted(spec.int_columns, key=lambda c: c.name), back.int_columns):
        assert a.name == b.name and np.array_equal(a.values, b.values) and np.array_equal(a.is_null, b.is_null)
    for a, b in zip(sorted(spec.dict_columns, key=lambda c: c.name), back.dict_columns):
        assert a.name == b.name and a.names == b.names and np.array_equal(a.is_null, b.is_null)
        valid = np.asarray(a.is_null) == 0
        assert np.array_equal(np.asarray(a.ids)[valid], np.asarray(b.ids)[valid])
        if a.lineage_parent is not None:
            assert np.array_equal(a.lineage_parent, b.lineage_parent)


def test_corrupted_archives_fail_cleanly(tmp_path):
    """Random byte changes and truncations of the saved table either still read (a changed value) or
    raise SiloFormatError: no other exception, no hang."""
    import random

    from silo_amd import silo_archive as sa

    d = ss.saved_state(tmp_path)
    with open(os.path.join(d, "default.silo"), "rb") as f:
        raw = f.read()
    with open(os.path.join(d, "database_schema.silo"), "rb") as f:
        schema = sa.read_schema(f.read())["default"]
    rng = random.Random(7)
    outcomes = {"read": 0, "refused": 0}
    for trial in range(300):
        data = bytearray(raw)
        if trial % 3 == 0:
            data = data[:rng.randrange(len(data))]
        else:
            for _ in range(rng.randrange(1, 4)):
                data[rng.randrange(len(data))] = rng.randrange(256)
        try:
            state = sa.read_table(bytes(data), schema)
            sa.table_spec(state)
            outcomes["read"] += 1
        except sa.SiloFormatError:
            outcomes["refused"] += 1
    assert outcomes["refused"] > 100


def test_round_trip_of_tables_without_containers_or_rows(tmp_path):
    """A column with no stored container (a null row, an all-missing row, a reference row), and a
    table with no rows, write and read back unchanged."""
    from silo_amd import silo_archive as sa, table
    from silo_amd.finalize import finalize_column

    for k, (rows, sizes) in enumerate(((["ACGT", None, "NNNN"], [3]), ([], []))):
        n = len(rows)
        spec = table.TableSpec(chunk_sizes=np.asarray(sizes, np.uint32))
        spec.seq_columns.append(finalize_column(ss.seq_column("main", 0, "ACGT", rows), spec.chunk_sizes))
        spec.int_columns.append(table.IntColumnData("d", True, np.arange(n, dtype=np.int32), np.zeros(n, np.uint8)))
        assert len(spec.seq_columns[0].key_position) == 0
        back = sa.load_database(sa.save_database(str(tmp_path / str(k)), {"default": spec}, 3)).table_spec()
        assert back.n_rows == n
        assert same(spec.seq_columns[0], back.seq_columns[0])
        if n:
            assert back.seq_columns[0].cov_end.tolist() == [4, 0, 0]
            assert back.seq_columns[0].is_null.tolist() == [0, 1, 0]
