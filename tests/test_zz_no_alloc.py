"""The allocation-free Mutations entry point (rdb_mutations_query_noalloc), called through the C ABI:
its rows equal rdb_mutations_query's on two calls in a row, the second on a smaller table (the shared
scratch keeps the first call's contents), with the dense and the sparse method; a table too large
for the scratch is rejected with RDB_EINVAL."""
import ctypes

import numpy as np
import pytest

from helpers import ExprGen, small_spec

pytestmark = pytest.mark.gpu


def call(et, fn, expr, column, min_proportion):
    from silo_amd.engine import _check, _p, lib

    cap, bufs = et._row_buffers(column)
    pos, fr, to, cnt, cov, prop = bufs
    n = ctypes.c_uint64()
    rc = getattr(lib(), fn)(et.handle, expr.encode(), column.encode(), float(min_proportion), _p(pos), _p(fr), _p(to),
                            _p(cnt), _p(cov), _p(prop), cap, ctypes.byref(n))
    _check(rc, et.engine.handle)
    k = n.value
    return [(int(pos[i]), int(fr[i]), int(to[i]), int(cnt[i]), int(cov[i]), float(prop[i])) for i in range(k)]


@pytest.fixture(scope="module")
def tables(engine):
    big_ds, big = small_spec(seed=61, n_rows=30000, L=600, chunk=4096)
    small_ds, small = small_spec(seed=62, n_rows=5000, L=300, chunk=2048)
    eb, es = engine.table(big), engine.table(small)
    yield (big_ds, eb), (small_ds, es)
    eb.close()
    es.close()


@pytest.mark.parametrize("sparse_max", [-1, 1 << 40])
def test_noalloc_matches_mutations_query(tables, engine, sparse_max):
    saved = engine.get_tuning("sparse_max_rows")
    engine.set_tuning("sparse_max_rows", sparse_max)
    try:
        for k, (ds, et) in enumerate(tables):  # the second call on the smaller table
            gen = ExprGen(ds, seed=70 + k)
            exprs = ["(true)", "(has_mut main 17)", f'(str_eq country "{ds.country_names[0]}")'] + \
                    [gen.expr(2) for _ in range(12)]
            for e in exprs:
                for p in (0.0, 0.05):
                    want = call(et, "rdb_mutations_query", e, "main", p)
                    assert call(et, "rdb_mutations_query_noalloc", e, "main", p) == want, (k, e, p)
    finally:
        engine.set_tuning("sparse_max_rows", saved)


def test_noalloc_rejects_a_table_too_large(engine):
    from silo_amd import table
    from silo_amd.engine import EngineError

    # 17,000 chunks of one row: the filter alone needs 8 KiB per chunk, more than 128 MiB
    n = 17000
    spec = table.TableSpec(chunk_sizes=np.ones(n, np.uint32))
    col = table.seq_column_from_strings("main", "ACGT", ["ACGT"] * n, 0)
    spec.seq_columns.append(col)
    et = engine.table(spec)
    try:
        with pytest.raises(EngineError) as err:
            call(et, "rdb_mutations_query_noalloc", "(true)", "main", 0.0)
        assert err.value.status == "RDB_EINVAL" and "scratch" in str(err.value)
        assert call(et, "rdb_mutations_query", "(true)", "main", 0.0) == []
    finally:
        et.close()
