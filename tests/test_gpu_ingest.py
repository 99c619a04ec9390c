"""Building sequence columns on the GPU.

The column builder (rdb_seq_builder_*) diffs aligned sequences with a HIP kernel: its diff equals the
oracle's restatement of extractCoverageAndMutationsFromSequence (aligned_sequence.cpp:18-122), with
offsets (SequenceColumnBuilder::insert, sequence_column.cpp:312-341), lower case, U, ambiguity
codes, gaps, empty and all-missing rows.  The local-reference adaptation and the sorted container
entries of every index build run on the GPU by default: the layout they give equals the host build
of the same rows byte for byte (the column digest), for the builder and for the diff-form upload,
and the tables answer like the oracle.  Illegal characters and over-long rows fail with the
reference's messages and leave the builder's rows unchanged."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NUC_REF = "ACGT"
NUC_ROW = "ACGTACGTRYKMSWBDHV-acgtuU"
AA_REF = "ACDEFGHIKLMNPQRSTVWY"
AA_ROW = "ACDEFGHIKLMNPQRSTVWY*-BZJOUacdx"


def random_rows(rng, ref: str, n: int, row_chars: str, missing: str, hot: list):
    """(sequence or None, offset) per row: nulls, empty rows, all-missing rows, rows with leading,
    trailing and internal missing runs and substitutions, rows placed at offsets.  At the `hot`
    positions most rows carry the same non-reference symbol, so adaptation flips them."""
    L = len(ref)
    rows = []
    for _ in range(n):
        u = rng.random()
        if u < 0.05:
            rows.append((None, 0))
            continue
        if u < 0.07:
            rows.append(("", int(rng.integers(0, L + 1))))
            continue
        if u < 0.09:
            rows.append((missing * L, 0))
            continue
        s = list(ref)
        for p in rng.choice(L, size=int(rng.integers(0, 12)), replace=False):
            s[p] = row_chars[int(rng.integers(0, len(row_chars)))]
        for p, c in hot:
            if rng.random() < 0.7:
                s[p] = c
        for _ in range(int(rng.integers(0, 4))):
            a = int(rng.integers(0, L))
            b = min(L, a + int(rng.integers(1, 9)))
            s[a:b] = missing * (b - a)
        if rng.random() < 0.3:
            k = int(rng.integers(1, 20))
            s[:k] = missing * k
        if rng.random() < 0.3:
            k = int(rng.integers(1, 20))
            s[L - k:] = missing * k
        seq = "".join(s)
        if rng.random() < 0.2:
            o = int(rng.integers(0, L))
            m = int(rng.integers(0, L - o + 1))
            rows.append((seq[o:o + m], o))
        else:
            rows.append((seq, 0))
    return rows


def oracle_column(name, alphabet, ref: str, rows):
    """The rows diffed by the oracle, in diff form: a row at offset o is its characters diffed
    against the reference's characters from o, positions shifted by o (aligned_sequence.cpp:28-35)."""
    from oracle import oracle
    from silo_amd import table

    n = len(rows)
    is_null = np.zeros(n, np.uint8)
    cs = np.zeros(n, np.uint32)
    ce = np.zeros(n, np.uint32)
    mptr = np.zeros(n + 1, np.uint64)
    nptr = np.zeros(n + 1, np.uint64)
    mpos, msym, ns, ne = [], [], [], []
    for i, (s, o) in enumerate(rows):
        if s is None:
            is_null[i] = 1
        else:
            st, en, missing, muts = oracle.diff_sequence(s, ref[o:o + len(s)], alphabet)
            if st < en:
                cs[i], ce[i] = st + o, en + o
            for p, sym in muts:
                mpos.append(p + o)
                msym.append(sym)
            missing = [p + o for p in missing]
            k = 0
            while k < len(missing):
                j = k
                while j + 1 < len(missing) and missing[j + 1] == missing[j] + 1:
                    j += 1
                ns.append(missing[k])
                ne.append(missing[j] + 1)
                k = j + 1
        mptr[i + 1] = len(mpos)
        nptr[i + 1] = len(ns)
    refsym = np.array([table.symbol_of(alphabet, c) for c in ref], np.uint8)
    return table.SeqColumnData(name, alphabet, refsym, is_null, cs, ce, mptr, np.array(mpos, np.uint32),
                               np.array(msym, np.uint8), nptr, np.array(ns, np.uint32), np.array(ne, np.uint32))


def aligned_column(name, alphabet, ref: str, rows, batch_rows: int):
    """The rows as builder batches of `batch_rows` rows (offsets given when any row has one)."""
    from silo_amd import table

    def batches():
        for a in range(0, len(rows), batch_rows):
            part = rows[a:a + batch_rows]
            chars = "".join(s or "" for s, _ in part).encode()
            lens = [len(s or "") for s, _ in part]
            row_ptr = np.zeros(len(part) + 1, np.uint64)
            row_ptr[1:] = np.cumsum(lens)
            offs = np.array([o for _, o in part], np.uint32)
            nulls = np.array([s is None for s, _ in part], np.uint8)
            yield np.frombuffer(chars, np.uint8), row_ptr, (offs if offs.any() else None), nulls

    refsym = np.array([table.symbol_of(alphabet, c) for c in ref], np.uint8)
    return table.AlignedSeqColumnData(name, alphabet, refsym, batches)


def meta_columns(n, rng):
    from silo_amd import table

    country = rng.integers(0, 3, n).astype(np.uint32)
    return [table.DictColumnData("country", country, np.zeros(n, np.uint8), ["CH", "DE", "FR"])]


def build(engine, sizes, seq_columns, rng, device: int):
    from silo_amd import table

    engine.set_tuning("index_build_device", device)
    try:
        spec = table.TableSpec(chunk_sizes=np.asarray(sizes, np.uint32), seq_columns=seq_columns,
                               dict_columns=meta_columns(int(np.sum(sizes)), rng))
        return engine.table(spec)
    finally:
        engine.set_tuning("index_build_device", 1)


def cases():
    rng = np.random.default_rng(31)
    nref = "".join(NUC_REF[i] for i in rng.integers(0, 4, 257))
    hot = [(p, "T" if nref[p] != "T" else "G") for p in (5, 64, 65, 128, 200)]
    yield "nucleotide", 0, nref, [700, 0, 300, 1000], random_rows(rng, nref, 2000, NUC_ROW, "N", hot), 333
    aref = "".join(AA_REF[i] for i in rng.integers(0, len(AA_REF), 131))
    hot = [(p, "W" if aref[p] != "W" else "Y") for p in (0, 63, 64, 130)]
    yield "amino_acid", 1, aref, [1500, 1200], random_rows(rng, aref, 2700, AA_ROW, "X", hot), 1000


@pytest.mark.parametrize("name,alphabet,ref,sizes,rows,batch_rows", list(cases()), ids=["nucleotide", "amino_acid"])
def test_builder_equals_host_build_and_oracle(engine, name, alphabet, ref, sizes, rows, batch_rows):
    from oracle.oracle import OracleTable
    from silo_amd import table

    rng = np.random.default_rng(5)
    diffed = oracle_column("s", alphabet, ref, rows)
    host = build(engine, sizes, [diffed], rng, device=0)
    tables = [host]
    try:
        tables.append(build(engine, sizes, [diffed], np.random.default_rng(5), device=1))
        tables.append(build(engine, sizes, [aligned_column("s", alphabet, ref, rows, batch_rows)],
                            np.random.default_rng(5), device=1))
        tables.append(build(engine, sizes, [aligned_column("s", alphabet, ref, rows, batch_rows)],
                            np.random.default_rng(5), device=0))
        want = host.digest("s")
        assert [t.digest("s") for t in tables] == [want] * 4
        spec = table.TableSpec(chunk_sizes=np.asarray(sizes, np.uint32), seq_columns=[diffed],
                               dict_columns=meta_columns(len(rows), np.random.default_rng(5)))
        ot = OracleTable(spec)
        try:
            local = ot.local_reference("s")
            assert not np.array_equal(local, diffed.reference)  # the hot positions flipped
            b = tables[2]
            assert np.array_equal(b.local_reference("s"), local)
            filters = ["(true)", "(has_mut s 5)", "(covered s 70)", "(not (covered s 0))", "(sym_eq s 64 -)"]
            for e in filters:
                assert np.array_equal(b.filter(e), ot.filter(e)), e
                assert b.mutations_query(e, "s", 0.0) == ot.mutations(e, "s", 0.0), e
                assert np.array_equal(b.mutation_counts(e, "s"), ot.mutation_counts(e, "s")), e
        finally:
            ot.close()
    finally:
        for t in tables:
            t.close()


def test_builder_on_generated_rows(engine):
    """Generated rows (nucleotide and two genes), rendered as aligned strings and built in batches:
    the same layout as the host build of their diff form."""
    from silo_amd import datagen, table

    ds = datagen.generate(datagen.GenSpec(n_rows=150_000, genome_length=3000, seed=41, lineage_model=True, n_nodes=60,
                                          n_mean=25, null_seq_frac=0.02), keep_handle=True)
    aa = datagen.translate(ds, datagen.gene_layout(ds.genome_length, [("g1", 400), ("g2", 300)], seed=3))
    spec = table.spec_from_dataset(ds, aa_columns=aa)
    host = None
    dev = None
    try:
        engine.set_tuning("index_build_device", 0)
        try:
            host = engine.table(spec)
        finally:
            engine.set_tuning("index_build_device", 1)
        aligned = table.TableSpec(chunk_sizes=spec.chunk_sizes, int_columns=spec.int_columns,
                                  dict_columns=spec.dict_columns)
        aligned.seq_columns = [table.AlignedSeqColumnData(c.name, c.alphabet, c.reference,
                                                          (lambda c=c: table.aligned_batches(c, 40_000)))
                               for c in spec.seq_columns]
        dev = engine.table(aligned)
        for c in spec.seq_columns:
            assert dev.digest(c.name) == host.digest(c.name), c.name
        for e in ["(true)", "(has_mut main 100)", "(covered g1 10)", "(and (has_mut g2 7) (covered main 2000))"]:
            assert dev.count(e) == host.count(e), e
    finally:
        for t in (host, dev):
            if t is not None:
                t.close()


def test_builder_errors(engine):
    """Illegal characters and over-long rows fail like the reference's insert; the failed append
    adds no rows, so appending the corrected rows afterwards builds the intended column."""
    import ctypes

    from silo_amd import engine as eng, table

    L = eng.lib()
    ctx = engine.handle
    EINVAL = 1

    def last_error():
        return L.rdb_last_error(ctx).decode()
    ref = "ACGTACGTAC"
    refsym = np.array([table.symbol_of(0, c) for c in ref], np.uint8)
    sizes = np.asarray([3], np.uint32)
    h = ctypes.c_void_p()
    assert L.rdb_table_create(ctx, sizes.ctypes.data_as(ctypes.c_void_p), 1, ctypes.byref(h)) == 0
    b = ctypes.c_void_p()
    try:
        assert L.rdb_seq_builder_create(h, b"s", 0, refsym.ctypes.data_as(ctypes.c_void_p), len(ref), ctypes.byref(b)) == 0

        def append(seqs, offsets=None):
            chars = np.frombuffer("".join(seqs).encode(), np.uint8).copy()
            ptr = np.zeros(len(seqs) + 1, np.uint64)
            ptr[1:] = np.cumsum([len(s) for s in seqs])
            offs = None if offsets is None else np.asarray(offsets, np.uint32)
            return L.rdb_seq_builder_append(b, chars.ctypes.data_as(ctypes.c_void_p), ptr.ctypes.data_as(ctypes.c_void_p),
                                            None if offs is None else offs.ctypes.data_as(ctypes.c_void_p), None,
                                            len(seqs))

        assert append(["ACGTACGTAC", "ACGTAZGTAC"]) == EINVAL
        assert last_error() == "illegal character 'Z' at position 5 in the input sequence"
        assert append(["ACG", "ACGT"], [7, 7]) == EINVAL
        assert last_error() == ("the sequence 'ACGT' which was inserted with an offset 7 is larger than the "
                                       "length of the reference genome: 10")
        assert append(["ACGTQ"], [2]) == EINVAL
        assert last_error() == "illegal character 'Q' at position 6 in the input sequence"
        assert L.rdb_seq_builder_finish(b) == EINVAL  # no rows yet
        assert append(["ACGTACGTAC", "NNGTTCGNNN"]) == 0
        assert append(["TT"], [4]) == 0
        assert L.rdb_seq_builder_finish(b) == 0
        assert L.rdb_seq_builder_finish(b) == EINVAL
        assert L.rdb_table_finalize(h) == 0
        loc = np.zeros(len(ref), np.uint8)
        assert L.rdb_table_local_reference(h, b"s", loc.ctypes.data_as(ctypes.c_void_p)) == 0
        # position 4: A, T, T over three rows -> T; position 5: C, C, T -> C
        assert "".join(table.NUC_CHARS[x] for x in loc) == "ACGTTCGTAC"
    finally:
        L.rdb_seq_builder_free(b)
        L.rdb_table_free(h)
