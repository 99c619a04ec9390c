"""Sorted date columns on the GPU: ranges on a null-free non-decreasing date32 column compile to
RangeSelection (date_between.cpp:75-79, 94-135) and evaluate as row ranges on the device, on the
single-query and the batched path, with an empty chunk, ranges ending exactly at chunk boundaries
and negations (range_selection.cpp:54-115); rows equal the oracle's."""
import numpy as np
import pytest

from helpers import sorted_date_filters, sorted_date_table

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def data(engine):
    from oracle.oracle import OracleTable

    spec, dates = sorted_date_table()
    ot = OracleTable(spec)
    et = engine.table(spec)
    yield spec, dates, ot, et
    et.close()
    ot.close()


def test_sorted_dates_compile_to_range_selection(data):
    spec, dates, ot, et = data
    assert et.explain("(date_between date 18100 18250)").startswith("RangeSelection")
    assert et.explain("(not (date_between date 18100 _))").startswith("RangeSelection")
    assert not et.explain("(date_between date_u 18100 18250)").startswith("RangeSelection")


def test_sorted_date_ranges_match_oracle(data):
    spec, dates, ot, et = data
    filters = sorted_date_filters(dates)
    want = [ot.filter(e) for e in filters]
    for e, w in zip(filters, want):
        assert np.array_equal(et.filter(e), w), e
    got = et.batch([{"filter": e, "action": "count"} for e in filters])
    assert got == [len(w) for w in want]
