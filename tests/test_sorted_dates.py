"""Sorted date columns (CPU only): a null-free non-decreasing date32 column takes the reference's
RangeSelection path (date_between.cpp:75-79, 94-135; range_selection.cpp:54-115), including an
empty chunk, ranges that end exactly at a chunk boundary and negated ranges.  The oracle agrees with
brute force."""
import numpy as np

from helpers import sorted_date_filters, sorted_date_table


def test_sorted_date_ranges_match_brute_force():
    from oracle import brute
    from oracle.oracle import OracleTable

    spec, dates = sorted_date_table()
    ot = OracleTable(spec)
    ints = {c.name: (c.values, c.is_null) for c in spec.int_columns}
    d = spec.dict_columns[0]
    bt = brute.BruteTable(spec.chunk_sizes, int_columns=ints, dict_columns={d.name: (d.ids, d.is_null, d.names, None)})
    try:
        assert ot.root_operator("(date_between date 18100 18250)") == "RangeSelection"
        assert ot.root_operator("(not (date_between date 18100 _))") == "RangeSelection"
        assert ot.root_operator("(date_between date_u 18100 18250)") == "Selection"
        for e in sorted_date_filters(dates):
            assert np.array_equal(ot.filter(e), bt.filter(e)), e
    finally:
        ot.close()
