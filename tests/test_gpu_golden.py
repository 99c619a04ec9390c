"""The engine (HIP, through the C ABI) reproduces the reference's own known-answer tests."""
import pytest

from kat_support import KATS, aggregate_names, dataset_spec, operator_spec, row_ids, symbol_chars

pytestmark = pytest.mark.gpu

STATUS = {"compile": "RDB_ECOMPILE", "query": "RDB_EQUERY"}


@pytest.mark.parametrize("case", KATS["operator_cases"], ids=lambda c: c["name"])
def test_operator_kats(engine, case):
    from silo_amd.engine import EngineError

    et = engine.table(operator_spec(case))
    try:
        if "expect_error" in case:
            with pytest.raises(EngineError) as err:
                et.filter(case["filter"])
            assert err.value.status == STATUS[case["expect_error"]]
        else:
            assert et.filter(case["filter"]).tolist() == case["expect"]
            assert et.count(case["filter"]) == len(case["expect"])
    finally:
        et.close()


@pytest.mark.parametrize("case", KATS["query_cases"], ids=lambda c: f'{c["dataset"]}-{c["name"]}')
def test_query_kats(engine, case):
    from silo_amd.engine import EngineError

    spec = dataset_spec(case["dataset"])
    et = engine.table(spec)
    try:
        if "expect_error" in case:
            with pytest.raises(EngineError) as err:
                et.count(case["filter"])
            assert err.value.status == STATUS[case["expect_error"]]
            return
        if "expect_rows" in case:
            assert et.filter(case["filter"]).tolist() == row_ids(spec, case["expect_rows"])
        if "expect_count" in case:
            assert et.count(case["filter"]) == case["expect_count"]
    finally:
        et.close()


@pytest.mark.parametrize("case", KATS["mutation_cases"], ids=lambda c: c["name"])
def test_mutation_kats(engine, case):
    et = engine.table(dataset_spec(case["dataset"]))
    try:
        chars = symbol_chars(0)
        rows = [[p, chars[f], chars[t], c, cov, prop] for p, f, t, c, cov, prop in
                et.mutations(case["filter"], case["column"], case["min_proportion"])]
        assert rows == case["expect"]
    finally:
        et.close()


@pytest.mark.parametrize("case", KATS["aggregate_cases"], ids=lambda c: c["name"])
def test_aggregate_kats(engine, case):
    from silo_amd.engine import EngineError

    spec = dataset_spec(case["dataset"])
    et = engine.table(spec)
    try:
        if "expect_error" in case:
            with pytest.raises(EngineError) as err:
                et.aggregate(case["filter"], case["dims"])
            assert err.value.status == STATUS[case["expect_error"]]
            return
        assert aggregate_names(spec, case["dims"], et.aggregate(case["filter"], case["dims"])) == case["expect"]
    finally:
        et.close()


@pytest.mark.parametrize("case", KATS["local_reference_cases"], ids=lambda c: c["dataset"])
def test_local_reference_kats(engine, case):
    et = engine.table(dataset_spec(case["dataset"]))
    try:
        assert "".join(symbol_chars(0)[s] for s in et.local_reference(case["column"])) == case["expect"]
    finally:
        et.close()


@pytest.mark.parametrize("mode", ["independence", "learned"])
def test_kat_counts_under_the_planner(engine, mode):
    """Planned evaluation (reordered children, short-circuit And / Or / n-of-m skips) of every
    operator and query known-answer filter gives the known cardinality."""
    mlp = engine.mlp(seed=13) if mode == "learned" else None
    try:
        engine.set_planner(mode, mlp)
        for case in KATS["operator_cases"]:
            if "expect_error" in case:
                continue
            et = engine.table(operator_spec(case))
            try:
                assert et.batch([{"filter": case["filter"], "action": "count"}]) == [len(case["expect"])], case["name"]
            finally:
                et.close()
        for case in KATS["query_cases"]:
            if "expect_error" in case:
                continue
            spec = dataset_spec(case["dataset"])
            et = engine.table(spec)
            try:
                want = case["expect_count"] if "expect_count" in case else len(case["expect_rows"])
                assert et.batch([{"filter": case["filter"], "action": "count"}]) == [want], case["name"]
            finally:
                et.close()
    finally:
        engine.set_planner("reference")
        if mlp is not None:
            mlp.free()
