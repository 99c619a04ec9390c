"""Bitset containers (CPU only): the table of `helpers.bitset_table` really holds stored symbols of
more than 4,096 scattered rows in a 65,536-row chunk (CRoaring keeps those as bitsets: an array
holds at most 4,096 values, roaring_container.cpp:10-13, and a run form only when it is smaller),
and the oracle's bitset branches agree with brute force over the materialized rows."""
import numpy as np
import pytest

from helpers import BITSET_FILTERS, bitset_table, brute_from_matrix


@pytest.fixture(scope="module")
def data():
    from oracle.oracle import OracleTable

    spec, M, null = bitset_table()
    ot = OracleTable(spec)
    bt = brute_from_matrix(spec, M, null)
    yield spec, M, null, ot, bt
    ot.close()


def test_matrix_diff_matches_oracle_restatement(data):
    from oracle.oracle import diff_sequence

    spec, M, null, ot, bt = data
    c = spec.seq_columns[0]
    ref = "".join("-ACGTRYSWKMBDHVN"[x] for x in c.reference)
    for i in list(range(0, len(M), 997)) + [5, 7, 13, 96, 1001]:
        if null[i]:
            continue
        s, e, miss, muts = diff_sequence("".join("-ACGTRYSWKMBDHVN"[x] for x in M[i]), ref)
        assert (s, e) == (int(c.cov_start[i]), int(c.cov_end[i])), i
        a, b = int(c.mut_ptr[i]), int(c.mut_ptr[i + 1])
        assert muts == list(zip(c.mut_pos[a:b].tolist(), c.mut_sym[a:b].tolist())), i
        a, b = int(c.nrun_ptr[i]), int(c.nrun_ptr[i + 1])
        assert miss == [p for x, y in zip(c.nrun_start[a:b], c.nrun_end[a:b]) for p in range(int(x), int(y))], i


def test_index_holds_bitset_shaped_containers(data):
    """Per full chunk, stored (non-local-reference) symbols at positions 10, 20, 31, 50 and 62 hold
    more than 4,096 rows in at least 2,048 runs (run form 4 bytes per run >= 8,192 bytes)."""
    spec, M, null, ot, bt = data
    local = ot.local_reference("main")
    assert local[50] == 1 and spec.seq_columns[0].reference[50] == 3  # G -> A: the reference symbol is stored
    found = []
    for p in (10, 20, 31, 50, 62):
        for c in range(2):
            rows = M[c * 65536:(c + 1) * 65536, p]
            keep = ~null[c * 65536:(c + 1) * 65536].astype(bool)
            for s in np.unique(rows):
                if s == local[p] or s == 15:
                    continue
                hit = (rows == s) & keep
                card = int(hit.sum())
                runs = int(np.sum(hit[1:] & ~hit[:-1]) + hit[0])
                if card > 4096 and 4 * runs >= 8192:
                    found.append((p, c, int(s)))
    assert {p for p, _, _ in found} == {10, 20, 31, 50, 62}, found


def test_oracle_filters_match_brute_force(data):
    spec, M, null, ot, bt = data
    for e in BITSET_FILTERS:
        assert np.array_equal(ot.filter(e), bt.filter(e)), e


def test_oracle_mutations_match_brute_force(data):
    spec, M, null, ot, bt = data
    for e in ["(true)", "(sym_eq main 10 T)", '(str_eq country "c1")', "(not (sym_eq main 31 -))",
              "(sym_eq main 50 C)"]:
        assert np.array_equal(ot.mutation_counts(e, "main"), bt.mutation_counts(e, "main")), e
        for p in (0.0, 0.05):
            assert ot.mutations(e, "main", p) == bt.mutations(e, "main", p), (e, p)


def test_oracle_aggregate_matches_brute_force(data):
    spec, M, null, ot, bt = data
    for e in ["(true)", "(has_mut main 20)"]:
        for dims in (["country"], ["main@50"], ["country", "main@10"]):
            assert ot.aggregate(e, dims) == bt.aggregate(e, dims), (e, dims)
