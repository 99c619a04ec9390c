"""The engine's host code under AddressSanitizer and UndefinedBehaviorSanitizer, with the device.

`sanitize_check device` (see test_sanitize.py) builds two tables from the same rows, one through
the device builder and one through the diff-form upload (their digests must agree), over chunks of
65,536, 3,000, 1 and 4,097 rows with bitset, array and run containers, a flipped local reference,
nulls and N runs, plus date, sorted date, integer, dictionary and lineage (with an alias) columns
and a literal bitmap.  It then runs every filter leaf kind through Count, Mutations, the
allocation-free entry point, the NDJSON and Arrow sinks and Aggregated, batches under each scan
mode against the single-query results, a submission whose second job fails, planned batches and
the estimator's entry points, and malformed uploads and queries, all with the host code
instrumented.  The results are checked for agreement between paths; their parity with the oracle is
checked by the other GPU tests."""
import pytest

from test_sanitize import run_checks

pytestmark = pytest.mark.gpu


def test_device_entry_points_under_sanitizers():
    print("sanitize_check device", flush=True)
    run_checks("device", 300)
