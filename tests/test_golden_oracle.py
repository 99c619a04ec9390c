"""The oracle reproduces the reference's own known-answer tests (CPU only)."""
import pytest

from kat_support import KATS, aggregate_names, dataset_spec, operator_spec, row_ids, symbol_chars


@pytest.mark.parametrize("case", KATS["operator_cases"], ids=lambda c: c["name"])
def test_operator_kats(case):
    from oracle.oracle import OracleError, OracleTable

    ot = OracleTable(operator_spec(case))
    if "expect_error" in case:
        with pytest.raises(OracleError):
            ot.filter(case["filter"])
    else:
        assert ot.filter(case["filter"]).tolist() == case["expect"]


@pytest.mark.parametrize("case", KATS["query_cases"], ids=lambda c: f'{c["dataset"]}-{c["name"]}')
def test_query_kats(case):
    from oracle.oracle import OracleError, OracleTable

    spec = dataset_spec(case["dataset"])
    ot = OracleTable(spec)
    if "expect_error" in case:
        with pytest.raises(OracleError):
            ot.count(case["filter"])
        return
    if "expect_rows" in case:
        assert ot.filter(case["filter"]).tolist() == row_ids(spec, case["expect_rows"])
    if "expect_count" in case:
        assert ot.count(case["filter"]) == case["expect_count"]


@pytest.mark.parametrize("case", KATS["mutation_cases"], ids=lambda c: c["name"])
def test_mutation_kats(case):
    from oracle.oracle import OracleTable

    spec = dataset_spec(case["dataset"])
    ot = OracleTable(spec)
    chars = symbol_chars(0)
    rows = [[p, chars[f], chars[t], c, cov, prop] for p, f, t, c, cov, prop in
            ot.mutations(case["filter"], case["column"], case["min_proportion"])]
    assert rows == case["expect"]


@pytest.mark.parametrize("case", KATS["aggregate_cases"], ids=lambda c: c["name"])
def test_aggregate_kats(case):
    from oracle.oracle import OracleError, OracleTable

    spec = dataset_spec(case["dataset"])
    ot = OracleTable(spec)
    if "expect_error" in case:
        with pytest.raises(OracleError):
            ot.aggregate(case["filter"], case["dims"])
        return
    assert aggregate_names(spec, case["dims"], ot.aggregate(case["filter"], case["dims"])) == case["expect"]


@pytest.mark.parametrize("case", KATS["local_reference_cases"], ids=lambda c: c["dataset"])
def test_local_reference_kats(case):
    from oracle.oracle import OracleTable

    spec = dataset_spec(case["dataset"])
    ot = OracleTable(spec)
    assert "".join(symbol_chars(0)[s] for s in ot.local_reference(case["column"])) == case["expect"]
