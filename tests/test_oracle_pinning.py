"""The oracle pinned against brute force over materialized sequences and against the generator's
own diff (CPU only)."""
import numpy as np
import pytest

from helpers import ExprGen, brute_for, random_bitmaps, small_spec


@pytest.fixture(scope="module", params=[(7, 1024, True), (23, 700, False)])
def data(request):
    from oracle.oracle import OracleTable

    seed, chunk, lineage = request.param
    ds, ts = small_spec(seed=seed, n_rows=2500, L=300, chunk=chunk, lineage=lineage)
    ts.bitmaps = random_bitmaps(ts, seed=seed)
    bt = brute_for(ds, ts)
    ot = OracleTable(ts)
    yield ds, ts, bt, ot
    ot.close()


def test_diff_restatement_matches_generator(data):
    from oracle.oracle import diff_sequence

    ds, ts, bt, ot = data
    ref = "".join("-ACGTRYSWKMBDHVN"[x] for x in ds.reference)
    for i in range(0, ds.n_rows, 7):
        if ds.is_null[i]:
            continue
        s, e, miss, muts = diff_sequence(ds.materialize(i), ref)
        assert (s, e) == (int(ds.cov_start[i]), int(ds.cov_end[i]))
        a, b = int(ds.mut_ptr[i]), int(ds.mut_ptr[i + 1])
        assert muts == list(zip(ds.mut_pos[a:b].tolist(), ds.mut_sym[a:b].tolist()))
        a, b = int(ds.nrun_ptr[i]), int(ds.nrun_ptr[i + 1])
        assert miss == [p for x, y in zip(ds.nrun_start[a:b], ds.nrun_end[a:b]) for p in range(int(x), int(y))]


def test_random_filters_match_brute_force(data):
    ds, ts, bt, ot = data
    gen = ExprGen(ds, seed=1, bitmaps=len(ts.bitmaps))
    for _ in range(200):
        e = gen.expr(3)
        assert np.array_equal(ot.filter(e), bt.filter(e)), e


def test_mutations_match_brute_force(data):
    ds, ts, bt, ot = data
    gen = ExprGen(ds, seed=2)
    for e in ["(true)", "(has_mut main 10)"] + [gen.expr(2) for _ in range(15)]:
        assert np.array_equal(ot.mutation_counts(e, "main"), bt.mutation_counts(e, "main")), e
        for p in (0.0, 0.05, 0.5, 1.0):
            assert ot.mutations(e, "main", p) == bt.mutations(e, "main", p), (e, p)


def test_aggregate_matches_brute_force(data):
    ds, ts, bt, ot = data
    gen = ExprGen(ds, seed=3)
    for e in ["(true)"] + [gen.expr(2) for _ in range(8)]:
        for dims in (["country"], ["lineage", "country"], ["main@33"], ["country", "main@5"]):
            assert ot.aggregate(e, dims) == bt.aggregate(e, dims), (e, dims)
