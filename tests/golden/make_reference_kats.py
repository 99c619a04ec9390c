"""Writes reference_kats.json: known-answer vectors transcribed from the reference's own unit and
query tests (inputs and expected outputs only).  Each case names the reference test it comes from.

Operator cases build the operator directly (op_threshold / op_intersection / op_union /
op_complement) over literal bitmaps in a row layout, as the reference's operator tests do with
IndexScan over roaring::Roaring literals and RowLayout::of(...).  Query cases run filter expressions
over small datasets given as aligned sequences and metadata, as the QUERY_TEST fixtures do; SaneQL
positions are 1-based, the expressions here are 0-based.
"""
import datetime
import json
import os

HERE = os.path.dirname(os.path.abspath(__file__))
OPS = "src/rhydb/query_engine/filter/operators"
SEX = "src/rhydb/query_engine/scalar_expressions"


def day(s: str) -> int:
    return (datetime.date.fromisoformat(s) - datetime.date(1970, 1, 1)).days


def th(n, exact, pos, neg):
    kids = [f"(bitmap {i})" for i in range(len(pos))] + [f"(not (bitmap {len(pos) + i}))" for i in range(len(neg))]
    return f"(op_threshold {n} {int(exact)} " + " ".join(kids) + ")"


def inter(pos, neg):
    kids = [f"(bitmap {i})" for i in range(len(pos))] + [f"(not (bitmap {len(pos) + i}))" for i in range(len(neg))]
    return "(op_intersection " + " ".join(kids) + ")"


operator_cases = []


def op_case(name, ref, rows, bitmaps, filt, expect=None, error=None):
    c = {"name": name, "ref": ref, "chunks": [rows] if rows is not None else [], "bitmaps": bitmaps, "filter": filt}
    if error:
        c["expect_error"] = error
    else:
        c["expect"] = expect
    operator_cases.append(c)


# threshold.test.cpp
T = f"{OPS}/threshold.test.cpp"
op_case("threshold_empty_children", f"{T}:28-39", 0, [], "(op_threshold 2 1)", error="compile")
neg = [[1, 2, 3], [1, 3]]
op_case("threshold_only_negated_1_exact", f"{T}:41-58", 4, neg, th(1, True, [], neg), [2])
op_case("threshold_only_negated_1_or_more", f"{T}:41-58", 4, neg, th(1, False, [], neg), [0, 2])
pos = [[1, 2], [1, 3], [1, 2, 3]]
for n, ex, want in [(1, True, []), (2, True, [2, 3]), (1, False, [1, 2, 3]), (2, False, [1, 2, 3])]:
    op_case(f"threshold_no_negated_{n}_{'exact' if ex else 'or_more'}", f"{T}:60-84", 4, pos, th(n, ex, pos, []), want)
pos = [[1, 2, 3], [1, 3], [1, 2, 3]]
neg = [[], [3]]
for n, ex, want in [(1, True, []), (2, True, [0]), (3, True, []), (4, True, [2, 3]), (1, False, [0, 1, 2, 3]),
                    (2, False, [0, 1, 2, 3]), (3, False, [1, 2, 3]), (4, False, [1, 2, 3])]:
    op_case(f"threshold_mixed_{n}_{'exact' if ex else 'or_more'}", f"{T}:86-162", 4, pos + neg, th(n, ex, pos, neg), want)
pos = [[1, 2, 3]]
neg = [[], [3], [4], [2, 4]]
for n, ex, want in [(1, True, []), (2, True, [4]), (3, True, []), (4, True, [0, 2, 3]), (1, False, [0, 1, 2, 3, 4]),
                    (2, False, [0, 1, 2, 3, 4]), (3, False, [0, 1, 2, 3]), (4, False, [0, 1, 2, 3])]:
    op_case(f"threshold_many_negated_{n}_{'exact' if ex else 'or_more'}", f"{T}:164-240", 5, pos + neg,
            th(n, ex, pos, neg), want)
# row layout of 4 rows while a negated child holds row 4: the DP's handling of that bit is pinned too
pos = [[]]
neg = [[3], [4], [2, 4]]
for n, ex, want in [(1, True, [4]), (2, True, [2, 3]), (3, True, [0, 1]), (1, False, [0, 1, 2, 3, 4]),
                    (2, False, [0, 1, 2, 3]), (3, False, [0, 1])]:
    op_case(f"threshold_empty_input_{n}_{'exact' if ex else 'or_more'}", f"{T}:242-300", 4, pos + neg,
            th(n, ex, pos, neg), want)
op_case("threshold_n_not_below_children", f"{T}:18-27 (constructor)", 4, [[1], [2]], th(2, False, [[1], [2]], []),
        error="compile")

# intersection.test.cpp
I = f"{OPS}/intersection.test.cpp"
op_case("intersection_without_non_negated", f"{I}:30-57", 5, [[1], [2]], inter([], [[1], [2]]), error="compile")
op_case("intersection_single_child", f"{I}:30-57", 5, [[1]], inter([[1]], []), error="compile")
pos = [[1, 2, 3], [1, 3], [1, 2, 3]]
op_case("intersection_no_negated", f"{I}:67-77", 5, pos, inter(pos, []), [1, 3])
op_case("intersection_with_negated", f"{I}:79-92", 5, pos + [[], [3]], inter(pos, [[], [3]]), [1])
op_case("intersection_many_negated", f"{I}:94-110", 5, [[1, 2, 3], [], [3], [4], [2, 4]],
        inter([[1, 2, 3]], [[], [3], [4], [2, 4]]), [1])
op_case("intersection_empty_input", f"{I}:112-125", 5, [[], [3], [4], [2, 4]], inter([[]], [[3], [4], [2, 4]]), [])

# union.test.cpp
U = f"{OPS}/union.test.cpp"


def uni(k):
    return "(op_union " + " ".join(f"(bitmap {i})" for i in range(k)) + ")"


op_case("union_empty_input", f"{U}:27-33", 5, [], uni(0), [])
op_case("union_one_input", f"{U}:35-42", 7, [[1, 3, 5]], uni(1), [1, 3, 5])
op_case("union_values_1", f"{U}:44-53", 7, [[1, 2, 3], [1, 3], [1, 2, 3]], uni(3), [1, 2, 3])
op_case("union_values_2", f"{U}:55-64", 8, [[1, 7], [1, 3], [3]], uni(3), [1, 3, 7])
op_case("union_many", f"{U}:66-87", 13, [[], [3], [4]] + [[2, 4]] * 11, uni(14), [2, 3, 4])
op_case("union_empty_bitmap", f"{U}:89-96", 80, [[]], uni(1), [])

# complement.test.cpp
C = f"{OPS}/complement.test.cpp"
op_case("complement_values", f"{C}:13-21", 5, [[1, 2, 3]], "(op_complement (bitmap 0))", [0, 4])
op_case("complement_empty_input", f"{C}:23-31", 3, [[]], "(op_complement (bitmap 0))", [0, 1, 2])
op_case("complement_empty_database", f"{C}:33-41", None, [[]], "(op_complement (bitmap 0))", [])
op_case("complement_full_input", f"{C}:43-51", 4, [[0, 1, 2, 3]], "(op_complement (bitmap 0))", [])
op_case("complement_single_input", f"{C}:53-61", 5, [[1]], "(op_complement (bitmap 0))", [0, 2, 3, 4])

# ---- datasets and query cases ------------------------------------------------------------------
datasets = {
    # scalar_expressions/symbol_equals.test.cpp:33-62 (nucleotide part)
    "symbol_equals": {"chunks": [5], "seq": {"segment1": [0, "ATGCN", ["ATGCN", "ATGCN", "NNNNN", "CATTT", None]]}},
    # symbol_equals.test.cpp (amino-acid part)
    "symbol_equals_aa": {"chunks": [4], "seq": {"gene1": [1, "M*", ["D*", "M*", "M*", "B*"]]}},
    # scalar_expressions/has_mutation.test.cpp:14-53
    "has_mutation": {"chunks": [4], "seq": {"segment1": [0, "ATGCN", ["ATGCN", "ATGCN", "NNNNN", "CATTT"]],
                                           "gene1": [1, "M*", ["M*", "C*", "M*", "X*"]]}},
    # scalar_expressions/nof.test.cpp:61-118
    "nof": {"chunks": [5],
            "dict": {"country": ["Switzerland", "Germany", "USA", "Switzerland", "France"],
                     "region": ["Europe", "Europe", "Americas", "Europe", "Europe"]},
            "date": {"date": [day("2020-01-01"), day("2021-06-15"), day("2019-03-20"), day("2018-11-01"),
                              day("2022-07-10")]}},
    # operators/union_all_node.test.cpp:15-47 (the Mutations known answer)
    "union_all": {"chunks": [4], "seq": {"segment1": [0, "A", ["T", "T", "T", "T"]]},
                  "dict": {"country": ["CH", "DE", "CH", "DE"]}},
    # filter/operators/is_in_covered_region.test.cpp:15-60: rows covering [0, 5) with the listed N
    # positions, written as aligned sequences (a trailing N is outside the covered range after the
    # diff, which does not change coverage at position 2)
    "covered": {"chunks": [8], "seq": {"s": [0, "AAAAA", ["ANNNA", "ANANA", "ANNNA", "AAAAA", "AAANA", "AAAAN",
                                                            "ANAAN", "AANAN"]]}},
    # storage/column/sequence_column.test.cpp:159-189
    "adapts_local_reference": {"chunks": [6], "seq": {"test_column": [0, "ACGT", ["ACTT", "ACTT", "ACTT", "ACNT",
                                                                                 "ACNT", "ACGT"]]}},
    # storage/column/sequence_column.test.cpp:191-223 (first finalize)
    "finalize_adapts": {"chunks": [4], "seq": {"test_column": [0, "ACGT", ["AAGT", "AAGT", "AAGT", "ACGT"]]}},
}

query_cases = []


def q(dataset, name, ref, filt, rows=None, count=None, error=None):
    c = {"dataset": dataset, "name": name, "ref": ref, "filter": filt}
    if rows is not None:
        c["expect_rows"] = rows
    if count is not None:
        c["expect_count"] = count
    if error is not None:
        c["expect_error"] = error
    query_cases.append(c)


SE = f"{SEX}/symbol_equals.test.cpp"
q("symbol_equals", "C_at_1", f"{SE}:75-79", "(sym_eq segment1 0 C)", count=1)
q("symbol_equals", "dot_at_1", f"{SE}:81-85", "(sym_eq segment1 0 .)", count=2)
q("symbol_equals", "C_at_2", f"{SE}:87-91", "(sym_eq segment1 1 C)", count=0)
q("symbol_equals", "C_at_3", f"{SE}:93-97", "(sym_eq segment1 2 C)", count=0)
for sym, counts in [("A", [2, 1, 0, 0, 0]), ("G", [0, 0, 2, 0, 0]), ("T", [0, 2, 1, 1, 1])]:
    for i, c in enumerate(counts):
        q("symbol_equals", f"{sym}_at_{i + 1}", SE, f"(sym_eq segment1 {i} {sym})", count=c)
q("symbol_equals", "N_at_1", SE, "(sym_eq segment1 0 N)", count=1)
q("symbol_equals", "N_at_5", SE, "(sym_eq segment1 4 N)", count=3)
q("symbol_equals", "out_of_range", SE, "(sym_eq segment1 999 C)", error="query")
q("symbol_equals", "out_of_range_low", SE, "(sym_eq segment1 -1 .)", error="query")
q("symbol_equals_aa", "D_at_1", SE, "(sym_eq gene1 0 D)", count=1)
q("symbol_equals_aa", "dot_at_1", SE, "(sym_eq gene1 0 .)", count=2)

HM = f"{SEX}/has_mutation.test.cpp"
q("has_mutation", "nucleotide_at_1", f"{HM}:55-61", "(has_mut segment1 0)", count=1)
q("has_mutation", "amino_acid_at_1", f"{HM}:63-68", "(has_mut gene1 0)", count=1)
q("has_mutation", "out_of_range", f"{HM}:70-76", "(has_mut segment1 1999)", error="query")
q("has_mutation", "out_of_range_edge_high", f"{HM}:86-92", "(has_mut segment1 5)", error="query")
q("has_mutation", "in_range_edge", f"{HM}:94-100", "(has_mut segment1 4)", count=1)

NO = f"{SEX}/nof.test.cpp"
CH, DE, US = '(str_eq country "Switzerland")', '(str_eq country "Germany")', '(str_eq country "USA")'
EU = '(str_eq region "Europe")'
D20 = f"(int_eq date {day('2020-01-01')})"
q("nof", "at_least_2_of_3", NO, f"(nof 2 0 {CH} {DE} {EU})", rows=[0, 1, 3])
q("nof", "at_least_3_of_3", NO, f"(nof 3 0 {CH} {D20} {EU})", rows=[0])
q("nof", "at_least_3_of_3_empty", NO, f"(nof 3 0 {CH} {DE} {EU})", rows=[])
q("nof", "at_least_1_of_2", NO, f"(nof 1 0 {CH} {US})", rows=[0, 2, 3])
q("nof", "exactly_1_of_2", NO, f"(nof 1 1 {CH} {EU})", rows=[1, 4])
q("nof", "exactly_2_of_2", NO, f"(nof 2 1 {CH} {EU})", rows=[0, 3])
q("nof", "at_least_0", NO, f"(nof 0 0 {CH} {DE})", rows=[0, 1, 2, 3, 4])
q("nof", "exactly_0", NO, f"(nof 0 1 {CH} {EU})", rows=[2])
q("nof", "with_negation", NO, f"(nof 2 0 {CH} (not {US}))", rows=[0, 3])
q("nof", "or_with_negated_child", NO, f"(nof 1 0 {CH} (not {DE}))", rows=[0, 2, 3, 4])
q("nof", "exactly_0_single_non_negated", NO, f"(nof 0 1 {CH})", rows=[1, 2, 4])
q("nof", "exactly_0_single_negated", NO, f"(nof 0 1 (not {CH}))", rows=[0, 3])
q("nof", "all_negated_and_case", NO, f"(nof 2 0 (not {CH}) (not {US}))", rows=[1, 4])
q("nof", "exactly_0_mixed_negated", NO, f"(nof 0 1 {CH} (not {US}))", rows=[2])
q("nof", "exact_count_exceeds_children", NO, f"(nof 5 1 {CH} {US})", rows=[])
q("nof", "at_least_count_exceeds_children", NO, f"(nof 5 0 {CH} {US})", rows=[])
q("nof", "single_negated_child", NO, f"(nof 1 0 (not {CH}))", rows=[1, 2, 4])
q("nof", "single_non_negated_child", NO, f"(nof 1 0 {CH})", rows=[0, 3])
q("nof", "exactly_1_of_3_threshold", NO, f"(nof 1 1 {CH} {DE} {EU})", rows=[4])
q("nof", "all_true_children", NO, "(nof 2 0 (true) (true) (true))", rows=[0, 1, 2, 3, 4])
q("nof", "all_true_exact_empty", NO, "(nof 1 1 (true) (true) (true))", rows=[])
q("nof", "all_false_children", NO, "(nof 2 0 (false) (false) (false))", rows=[])
q("nof", "mixed_trivial_and_real", NO, f"(nof 1 0 (true) (false) {CH})", rows=[0, 1, 2, 3, 4])
q("nof", "negative_count", NO, f"(nof -1 0 {CH})", rows=[0, 1, 2, 3, 4])

IC = f"{OPS}/is_in_covered_region.test.cpp"
q("covered", "covered_at_2", f"{IC}:15-60", "(covered s 2)", rows=[1, 3, 4, 5, 6])
q("covered", "not_covered_at_2", f"{IC}:15-60", "(not (covered s 2))", rows=[0, 2, 7])

mutation_cases = [
    {"dataset": "union_all", "name": "mutations_country_CH", "ref": "src/rhydb/query_engine/operators/"
     "union_all_node.test.cpp:185-194", "filter": '(str_eq country "CH")', "column": "segment1",
     "min_proportion": 0.0, "expect": [[1, "A", "T", 2, 2, 1.0]]},
    {"dataset": "union_all", "name": "mutations_country_DE", "ref": "src/rhydb/query_engine/operators/"
     "union_all_node.test.cpp:185-194", "filter": '(str_eq country "DE")', "column": "segment1",
     "min_proportion": 0.0, "expect": [[1, "A", "T", 2, 2, 1.0]]},
]

local_reference_cases = [
    {"dataset": "adapts_local_reference", "column": "test_column", "ref": "src/rhydb/storage/column/"
     "sequence_column.test.cpp:159-189", "expect": "ACTT"},
    {"dataset": "finalize_adapts", "column": "test_column", "ref": "src/rhydb/storage/column/"
     "sequence_column.test.cpp:191-212", "expect": "AAGT"},
    {"dataset": "symbol_equals", "column": "segment1", "ref": "derived: position 5 has reference N and one covered "
     "row with T", "expect": "ATGCT"},
]

with open(os.path.join(HERE, "reference_kats.json"), "w") as f:
    json.dump({"operator_cases": operator_cases, "datasets": datasets, "query_cases": query_cases,
               "mutation_cases": mutation_cases, "local_reference_cases": local_reference_cases}, f, indent=1)
print(f"{len(operator_cases)} operator cases, {len(query_cases)} query cases, {len(mutation_cases)} mutation cases")
