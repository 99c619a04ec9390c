"""Result sinks of the Mutations action (host only, no GPU): NDJSON lines as the reference's
NdjsonSink writes them (exec_node/ndjson_sink.cpp: one object per row, fields in
MutationsNode::getOutputSchema order, strings JSON-escaped, doubles in shortest round-trip form:
the form Python's json module writes) and an Arrow IPC stream as its ArrowIpcSink writes it
(exec_node/arrow_ipc_sink.cpp, read back here with pyarrow), for every field subset."""
import json

import numpy as np
import pytest

ORDER = ["mutationFrom", "mutationTo", "sequenceName", "position", "proportion", "coverage", "count"]
NUC = "-ACGTRYSWKMBDHVN"
AA = "-ACDEFGHIKLMNOPQRSTUVWYBJZ*X"


def random_rows(seed, n, alphabet=0):
    rng = np.random.default_rng(seed)
    nsym = 16 if alphabet == 0 else 28
    rows = []
    for _ in range(n):
        cov = int(rng.integers(1, 4_000_000))
        cnt = int(rng.integers(1, cov + 1))
        rows.append((int(rng.integers(1, 30000)), int(rng.integers(0, nsym)), int(rng.integers(0, nsym)), cnt, cov,
                     cnt / cov))
    rows += [(5, 1, 2, 1, 100000, 1e-05), (6, 1, 2, 1, 10000, 1e-04), (7, 2, 3, 3, 3, 1.0), (8, 3, 4, 1, 10, 0.1),
             (9, 4, 0, 2, 3, 2 / 3)]
    return rows


def expected_ndjson(rows, name, fields, alphabet=0):
    chars = NUC if alphabet == 0 else AA
    out = []
    for pos, fr, to, cnt, cov, prop in rows:
        vals = {"mutationFrom": chars[fr], "mutationTo": chars[to], "sequenceName": name, "position": pos,
                "proportion": prop, "coverage": cov, "count": cnt}
        obj = {k: vals[k] for i, k in enumerate(ORDER) if fields & (1 << i)}
        out.append(json.dumps(obj, separators=(",", ":"), ensure_ascii=False) + "\n")
    return "".join(out)


FIELD_SETS = [127, 1, 8 | 16, 2 | 4 | 64, 127 & ~4, 16]


@pytest.mark.parametrize("fields", FIELD_SETS)
@pytest.mark.parametrize("alphabet", [0, 1])
def test_ndjson_matches_json_rendering(fields, alphabet):
    from silo_amd.engine import mutation_rows_ndjson

    rows = random_rows(fields + alphabet, 300, alphabet)
    name = "main" if alphabet == 0 else "ORF1a"
    assert mutation_rows_ndjson(rows, name, alphabet, fields) == expected_ndjson(rows, name, fields, alphabet)


def test_ndjson_escapes_names_and_formats_doubles():
    from silo_amd.engine import mutation_rows_ndjson

    name = 'seq "1"\\\t\n\x01é/'
    rows = random_rows(3, 20)
    assert mutation_rows_ndjson(rows, name) == expected_ndjson(rows, name, 127)
    # shortest round-trip doubles over the proportions count / coverage can take, and edge values
    rng = np.random.default_rng(5)
    cov = rng.integers(1, 20_000_000, 20000)
    props = [int(c) / int(v) for c, v in zip(rng.integers(1, cov + 1), cov)] + [1e-7, 1.5e-5, 9.99999e-5, 0.25]
    rows = [(1, 1, 2, 1, 1, p) for p in props]
    got = mutation_rows_ndjson(rows, "m", 0, 16).splitlines()
    assert got == ['{"proportion":' + repr(p) + "}" for p in props]


def test_empty_and_invalid():
    from silo_amd.engine import EngineError, mutation_rows_arrow_ipc, mutation_rows_ndjson

    assert mutation_rows_ndjson([], "main") == ""
    with pytest.raises(EngineError) as e:
        mutation_rows_ndjson([(1, 1, 2, 1, 1, 1.0)], "main", 0, 0)
    assert e.value.status == "RDB_EINVAL"
    with pytest.raises(EngineError):
        mutation_rows_ndjson([(1, 1, 2, 1, 1, 1.0)], "main", 0, 128)
    with pytest.raises(EngineError):
        mutation_rows_arrow_ipc([(1, 1, 16, 1, 1, 1.0)], "main", 0)


@pytest.mark.parametrize("fields", FIELD_SETS)
@pytest.mark.parametrize("n", [0, 1, 257])
def test_arrow_ipc_stream_reads_back(fields, n):
    pa = pytest.importorskip("pyarrow")
    import pyarrow.ipc as ipc

    from silo_amd.engine import mutation_rows_arrow_ipc

    rows = random_rows(n + fields, n)[:n]
    data = mutation_rows_arrow_ipc(rows, "main", 0, fields)
    reader = ipc.open_stream(pa.BufferReader(data))
    table = reader.read_all()
    names = [k for i, k in enumerate(ORDER) if fields & (1 << i)]
    types = {"mutationFrom": pa.string(), "mutationTo": pa.string(), "sequenceName": pa.string(),
             "position": pa.int32(), "proportion": pa.float64(), "coverage": pa.int32(), "count": pa.int32()}
    assert table.schema.names == names
    assert [f.type for f in table.schema] == [types[k] for k in names]
    assert all(f.nullable for f in table.schema)
    assert table.num_rows == n
    cols = table.to_pydict()
    for i, (pos, fr, to, cnt, cov, prop) in enumerate(rows):
        vals = {"mutationFrom": NUC[fr], "mutationTo": NUC[to], "sequenceName": "main", "position": pos,
                "proportion": prop, "coverage": cov, "count": cnt}
        for k in names:
            assert cols[k][i] == vals[k], (i, k)
