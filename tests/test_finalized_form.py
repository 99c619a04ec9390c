"""The finalized storage of a sequence column (CPU only): the oracle's export of the state a
finalized reference SequenceColumn holds (local reference, vertical containers with CRoaring
typecodes and container_write bytes, covered ranges, missing positions, nulls) reads back into
exactly the diff form the column was ingested from, on tables with adapted local references,
bitset / array / run containers, missing runs, null rows and an amino-acid column."""
import numpy as np
import pytest

from finalized_support import to_diff_form
from helpers import bitset_table, small_spec


def tables():
    from silo_amd import datagen, table

    ds, ts = small_spec(seed=11, n_rows=3000, L=400, chunk=1024)
    yield "lineage", ts
    spec, _, _ = bitset_table()
    yield "bitset", spec
    ds = datagen.generate(datagen.GenSpec(n_rows=6000, genome_length=900, seed=77, lineage_model=True, n_nodes=40,
                                          n_mean=40, null_seq_frac=0.02), keep_handle=True)
    aa = datagen.translate(ds, datagen.gene_layout(ds.genome_length, [("g1", 120), ("g2", 150)], seed=5))
    yield "amino_acid", table.spec_from_dataset(ds, chunk=2048, aa_columns=aa)


@pytest.mark.parametrize("name,spec", list(tables()), ids=lambda x: x if isinstance(x, str) else "")
def test_finalized_form_reads_back_into_the_diff_form(name, spec):
    from oracle.oracle import OracleTable

    ot = OracleTable(spec)
    try:
        for c in spec.seq_columns:
            f = ot.export_finalized(c.name)
            assert set(np.unique(f.typecode).tolist()) <= {1, 2, 3}
            if name == "bitset":
                assert {1, 2, 3} <= set(np.unique(f.typecode).tolist())
            assert np.array_equal(f.local_reference, ot.local_reference(c.name))
            assert np.array_equal(f.is_null, c.is_null)
            keep = c.is_null == 0
            assert np.array_equal(f.cov_start[keep], np.asarray(c.cov_start)[keep])
            assert np.array_equal(f.cov_end[keep], np.asarray(c.cov_end)[keep])
            mut_ptr, pos, sym, nrun_ptr, ns, ne = to_diff_form(f, spec.chunk_sizes)
            assert np.array_equal(mut_ptr, c.mut_ptr)
            assert np.array_equal(pos, c.mut_pos) and np.array_equal(sym, c.mut_sym)
            assert np.array_equal(nrun_ptr, c.nrun_ptr)
            assert np.array_equal(ns, c.nrun_start) and np.array_equal(ne, c.nrun_end)
            if name == "lineage":
                assert (f.local_reference != f.reference).any()
    finally:
        ot.close()


def decoded_index(f) -> dict:
    from silo_amd import silo_archive as sa

    out = {}
    for k in range(len(f.key_position)):
        o = int(f.payload_offset[k])
        rows = sa.container_values(int(f.typecode[k]), int(f.cardinality[k]), bytes(f.payload[o:o + 8194]))
        assert len(rows) == int(f.cardinality[k])
        out[(int(f.key_position[k]), int(f.key_chunk[k]), int(f.key_symbol[k]))] = rows.tolist()
    return out


@pytest.mark.parametrize("name,spec", list(tables()), ids=lambda x: x if isinstance(x, str) else "")
def test_host_finalize_equals_the_oracle_export(name, spec):
    """silo_amd.finalize (numpy, used to write saved states) gives the oracle's finalized column:
    the same local reference, containers with the same rows (and typecodes, both run-optimized),
    covered ranges, missing positions and nulls."""
    from oracle.oracle import OracleTable
    from silo_amd.finalize import finalize_column

    ot = OracleTable(spec)
    try:
        for c in spec.seq_columns:
            want = ot.export_finalized(c.name)
            got = finalize_column(c, spec.chunk_sizes)
            assert got.local_reference.tolist() == want.local_reference.tolist()
            assert decoded_index(got) == decoded_index(want)
            assert got.typecode.tolist() == want.typecode.tolist()
            for a in ("is_null", "cov_start", "cov_end"):
                assert np.asarray(getattr(got, a)).tolist() == np.asarray(getattr(want, a)).tolist(), a
            for x, y in zip(got.missing_positions(), want.missing_positions()):
                assert x.tolist() == y.tolist()
            for x, y in zip(got.missing_runs(), want.missing_runs()):
                assert x.tolist() == y.tolist()
    finally:
        ot.close()
