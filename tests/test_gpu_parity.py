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
