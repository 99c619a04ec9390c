"""The Mutations action written to the result sinks on the GPU (rdb_mutations_ndjson,
rdb_mutations_arrow_ipc): the NDJSON text and the Arrow record batch hold exactly the oracle's
Mutations rows (mutations_node.cpp:292-366), on a nucleotide and an amino-acid column."""
import pytest

from helpers import ExprGen
from test_sinks import ORDER, expected_ndjson

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def data(engine):
    from oracle.oracle import OracleTable
    from silo_amd import datagen, table

    ds = datagen.generate(datagen.GenSpec(n_rows=8000, genome_length=900, seed=91, lineage_model=True, n_nodes=40,
                                          n_mean=40, null_seq_frac=0.02), keep_handle=True)
    aa = datagen.translate(ds, datagen.gene_layout(ds.genome_length, [("g1", 120), ("g2", 150)], seed=5))
    spec = table.spec_from_dataset(ds, chunk=4096, aa_columns=aa)
    ot = OracleTable(spec)
    et = engine.table(spec)
    yield ds, ot, et
    et.close()
    ot.close()


def test_sinks_hold_the_oracle_rows(data):
    pa = pytest.importorskip("pyarrow")
    import pyarrow.ipc as ipc

    ds, ot, et = data
    gen = ExprGen(ds, seed=12)
    for e in ["(true)", "(has_mut main 40)"] + [gen.expr(2) for _ in range(8)]:
        for col, alphabet in (("main", 0), ("g1", 1)):
            for fields in (127, 8 | 16 | 64):
                want = ot.mutations(e, col, 0.05)
                assert et.mutations_ndjson(e, col, 0.05, fields) == expected_ndjson(want, col, fields, alphabet), (e, col)
            table = ipc.open_stream(pa.BufferReader(et.mutations_arrow_ipc(e, col, 0.05))).read_all()
            assert table.schema.names == ORDER
            cols = table.to_pydict()
            chars = "-ACGTRYSWKMBDHVN" if alphabet == 0 else "-ACDEFGHIKLMNOPQRSTUVWYBJZ*X"
            got = [(p, chars.index(f), chars.index(t), c, v, pr) for p, f, t, c, v, pr in
                   zip(cols["position"], cols["mutationFrom"], cols["mutationTo"], cols["count"], cols["coverage"],
                       cols["proportion"])]
            assert got == want, (e, col)
            assert set(cols["sequenceName"]) <= {col}
