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
    """rows: rows of a one-chunk layout, a list of chunk sizes, or None for the empty layout."""
    chunks = [] if rows is None else (list(rows) if isinstance(rows, list) else [rows])
    c = {"name": name, "ref": ref, "chunks": chunks, "bitmaps": bitmaps, "filter": filt}
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

# range_selection.test.cpp: RangeSelection(ranges, RowLayout::of(...)) and RangeSelection::negate
R = f"{OPS}/range_selection.test.cpp"


def rng(*pairs):
    return "(op_range " + " ".join(str(x) for x in pairs) + ")"


for name, lines, layout, ranges, want, want_neg in [
        ("values", "10-22", 8, (0, 2, 3, 5), [0, 1, 3, 4], [2, 5, 6, 7]),
        ("empty_database", "24-32", None, (), [], []),
        ("empty_ranges", "34-45", 9, (0, 0, 4, 4), [], list(range(9))),
        ("full_range", "47-59", 8, (0, 1 << 16), list(range(8)), []),
        ("meeting_ranges", "61-72", 9, (0, 2, 2, 4), [0, 1, 2, 3], [4, 5, 6, 7, 8]),
        ("spanning_chunks", "74-106", [4, 3, 5], (2, (2 << 16) | 2),
         [2, 3, 1 << 16, (1 << 16) + 1, (1 << 16) + 2, 2 << 16, (2 << 16) + 1],
         [0, 1, (2 << 16) + 2, (2 << 16) + 3, (2 << 16) + 4])]:
    op_case(f"range_selection_{name}", f"{R}:{lines}", layout, [], rng(*ranges), want)
    op_case(f"range_selection_{name}_negated", f"{R}:{lines}", layout, [], f"(not {rng(*ranges)})", want_neg)

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

# filter/operators/selection.test.cpp: a single int32 chunk (makeTestColumn)
SELECTION_VALUES = [0, 1, 4, 4, 4, 1, 1, 1, 1, 1]
datasets["selection"] = {"chunks": [10], "int": {"test": SELECTION_VALUES}}
datasets["selection_negative"] = {"chunks": [10], "int": {"test": [0, -1, 4, 4, 4, -1, -1, -1, -1, -1]}}
# makeRangeTestColumn: value == row index over 100 rows; literal child bitmaps of the BitmapProducer
datasets["selection_range"] = {"chunks": [100], "int": {"test": list(range(100))},
                               "bitmaps": [list(range(5, 80)), [5, 15, 45, 60, 99]]}
# scalar_expressions/date_between.test.cpp: one row; both date columns hold a single value (sorted)
datasets["date_between"] = {"chunks": [1], "date": {"sorted_date": [day("2020-12-24")],
                                                   "unsorted_date": [day("2023-01-20")]},
                            "dict": {"primaryKey": ["id"]}}
# scalar_expressions/lineage_filter.test.cpp: RECOMBINANT has two parents; recombinant edges are not
# followed in the default mode, so here it is a root
datasets["lineage_filter"] = {"chunks": [5], "dict": {"pango_lineage": ["BASE.1", "BASE.1", "CHILD", None, "RECOMBINANT"]},
                              "lineage_parents": {"pango_lineage": {"CHILD": "BASE.1", "CHILD.2": "BASE.1"}},
                              "extra_values": {"pango_lineage": ["CHILD.2"]}}
# scalar_expressions/or.test.cpp:409-461
datasets["or"] = {"chunks": [6], "dict": {"country": ["Switzerland", "Germany", "USA", "Switzerland", "France", "Germany"],
                                          "region": ["Europe", "Europe", "Americas", "Europe", "Europe", "Europe"]}}
# operators/bitmap_aggregation_node.test.cpp:16-70 (reference ATGCN / M*)
datasets["bitmap_aggregation"] = {"chunks": [4], "seq": {"segment1": [0, "ATGCN", ["ATGCN", "ATGCN", "NNNNN", "CATTT"]],
                                                         "gene1": [1, "M*", ["M*", "C*", "M*", "X*"]]},
                                  "dict": {"region": ["Europe", "Europe", "Asia", "Europe"]}}
datasets["bitmap_aggregation_nulls"] = {"chunks": [4], "seq": {"segment1": [0, "ATGCN", ["ATGCN", "ATGCN", None, "CATTT"]],
                                                               "gene1": [1, "M*", ["M*", "M*", "M*", None]]},
                                        "dict": {"region": ["Europe"] * 4}}
datasets["bitmap_aggregation_ambiguous"] = {"chunks": [4], "seq": {"segment1": [0, "ATGCN", ["ATGCN", "RTGCN", "RTGCN",
                                                                                                 "YTGCN"]]},
                                            "dict": {"region": ["Europe"] * 4}}
# storage/column/vertical_sequence_index.test.cpp:311-441: one position; N rows are the uncovered
# ones of each test's coverage list
VSI = "src/rhydb/storage/column/vertical_sequence_index.test.cpp"
vsi_cases = [
    ("adapts_single_symbol", "311-335", "C", ["A"], "A", {"A": [0], "C": []}),
    ("adapts_single_symbol_with_missing", "337-357", "C", ["A", "N", "N"], "A", {"A": [0], "C": []}),
    ("does_not_adapt_single_symbol", "359-371", "C", ["A", "C"], "C", {"A": [0], "C": [1]}),
    ("does_not_adapt_with_missing", "373-386", "T", ["A", "N", "T", "N", "T", "T", "A"], "T",
     {"A": [0, 6], "T": [2, 4, 5]}),
    ("adapts_and_flips", "388-410", "T", ["A", "N", "T", "N", "A", "T", "A"], "A", {"A": [0, 4, 6], "T": [2, 5]}),
    ("adapts_and_flips_many_symbols", "412-441", "T", ["A", "N", "T", "N", "A", "-", "-", "-", "T", "A"], "-",
     {"A": [0, 4, 9], "T": [2, 8], "-": [5, 6, 7]}),
]
for name, _, ref, rows, _, _ in vsi_cases:
    datasets[f"vsi_{name}"] = {"chunks": [len(rows)], "seq": {"s": [0, ref, rows]}}
# storage/column/horizontal_coverage_index.test.cpp: GENOME_LENGTH 100, sequences inserted at an
# offset (written here as aligned rows padded with N); None is a null sequence
HCI = "src/rhydb/storage/column/horizontal_coverage_index.test.cpp"
HL = 100


hci_cases = [
    # (name, lines, [(sequence, offset) | None], {position: covered rows})
    ("null_sequence", "57-63", [None], {0: [], 50: [], 99: []}),
    ("multiple_nulls", "65-72", [None, None, None], {0: [], 50: []}),
    ("covered_positions", "99-107", [("ACGT", 0)], {0: [0], 1: [0], 2: [0], 3: [0]}),
    ("uncovered_positions", "109-116", [("ACGT", 0)], {4: [], 10: [], 50: []}),
    ("overlap", "118-137", [("ACGTACGT", 0), ("TGCA", 5), ("AAAA", 10)],
     {6: [0, 1], 5: [0, 1], 7: [0, 1], 8: [1], 11: [2]}),
    ("non_overlapping", "139-151", [("ACGT", 0), ("TGCA", 10)], {2: [0], 11: [1], 5: []}),
    ("n_is_uncovered", "154-165", [("ACNGT", 0)], {0: [0], 1: [0], 3: [0], 4: [0], 2: []}),
    ("multiple_ns", "167-180", [("ANNNGTA", 0)], {0: [0], 4: [0], 5: [0], 6: [0], 1: [], 2: [], 3: []}),
    ("all_ns", "182-193", [("NNNN", 5)], {5: [], 6: [], 7: [], 8: [], 0: []}),
    ("n_at_same_position", "195-205", [("ACNGT", 0), ("AANAT", 0), ("ACGGT", 0)], {2: [2], 0: [0, 1, 2]}),
    ("n_and_offset", "207-218", [("ACNGT", 10)], {12: [], 10: [0], 11: [0], 13: [0], 14: [0]}),
    ("lowercase_n", "225-230", [("ACnGT", 0)], {2: []}),
    ("genome_boundary", "238-242", [("ACGT", 0)], {HL - 1: []}),
    ("at_end_of_genome", "253-263", [("ACG", HL - 3)], {HL - 3: [0], HL - 2: [0], HL - 1: [0], HL - 4: []}),
    ("n_at_end_of_genome", "265-274", [("ACN", HL - 3)], {HL - 3: [0], HL - 2: [0], HL - 1: []}),
    ("mixed_null_and_covered", "282-295", [("ACGT", 0), None, ("TGCA", 0)],
     {0: [0, 2], 1: [0, 2], 2: [0, 2], 3: [0, 2], 10: []}),
    ("nulls_with_coverage", "297-309", [None, None, ("ACGT", 5), None], {0: [], 5: [2], 6: [2]}),
    ("null_coverage_and_ns", "311-321", [("ACNGT", 0), None, ("ACGGT", 0)], {2: [2], 0: [0, 2]}),
    ("completely_overlapping", "324-337", [("AAAA", 5), ("TTTT", 5), ("GGGG", 5)],
     {5: [0, 1, 2], 6: [0, 1, 2], 7: [0, 1, 2], 8: [0, 1, 2], 4: []}),
    ("overlapping_with_n", "339-355", [("ANNN", 5), ("NANN", 5), ("NNAN", 5)], {5: [0], 6: [1], 7: [2], 8: []}),
    ("partially_overlapping", "357-373", [("AAAAAAAA", 0), ("TTTTTTTT", 4), ("GGGGGGGG", 8)],
     {2: [0], 5: [0, 1], 9: [1, 2], 14: [2]}),
    ("adjacent", "375-384", [("AAAA", 0), ("TTTT", 4), ("GGGG", 8)], {3: [0], 4: [1], 7: [1], 8: [2]}),
    ("many_sequences", "387-401", [("ACGT", 0)] * 1000, {0: list(range(1000)), 10: []}),
    ("many_at_different_positions", "403-414", [("A", i) for i in range(100)], {p: [p] for p in range(100)}),
    ("long_sequence", "416-429", [("A" * (HL // 2), 0)], {0: [0], HL // 4: [0], HL // 2 - 1: [0], HL // 2: []}),
    ("in_order", "432-442", [("A", 0), ("T", 0), ("G", 0)], {0: [0, 1, 2], 5: []}),
    ("out_of_order", "444-454", [("G", 0), ("A", 0), ("T", 0)], {0: [0, 1, 2], 5: []}),
    ("empty_sequence", "457-464", [("", 5)], {0: [], 5: [], 10: []}),
]
for name, _, rows, _ in hci_cases:
    # "placed": rows are [sequence, offset] (or null) inside a genome of the reference's length, the
    # positions outside the sequence missing, as insertSequenceCoverage(sequence, offset) places them
    datasets[f"hci_{name}"] = {"chunks": [len(rows)], "placed": {"s": [0, "A" * HL, [None if r is None else list(r)
                                                                                   for r in rows]]}}

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

SEL = f"{OPS}/selection.test.cpp"
for cmp, lines, want in [("eq", "38-51", [1, 5, 6, 7, 8, 9]), ("ne", "53-68", [0, 2, 3, 4]), ("lt", "70-83", [0]),
                         ("le", "85-100", [0, 1, 5, 6, 7, 8, 9]), ("ge", "102-117", list(range(1, 10))),
                         ("gt", "119-132", [2, 3, 4])]:
    q("selection", f"compare_{cmp}", f"{SEL}:{lines}", f"(op_select (cmp test {cmp} 1))", rows=want)
    q("selection", f"compare_{cmp}_negated", f"{SEL}:{lines}", f"(not (op_select (cmp test {cmp} 1)))",
      rows=[i for i in range(10) if i not in want])
q("selection_negative", "negative_numbers", f"{SEL}:134-145", "(op_select (cmp test eq -1))", rows=[1, 5, 6, 7, 8, 9])
q("selection_range", "two_predicates_without_child", f"{SEL}:192-199",
  "(op_select (cmp test ge 10) (cmp test lt 50))", rows=list(range(10, 50)))
q("selection_range", "two_predicates_large_child", f"{SEL}:201-216",
  "(op_select (bitmap 0) (cmp test ge 10) (cmp test lt 50))", rows=list(range(10, 50)))
q("selection_range", "two_predicates_small_child", f"{SEL}:218-232",
  "(op_select (bitmap 1) (cmp test ge 10) (cmp test lt 50))", rows=[15, 45])
q("selection_range", "no_predicates", f"{OPS}/selection.cpp (constructor)", "(op_select (bitmap 1))", error="compile")

DB = f"{SEX}/date_between.test.cpp"
d = day("2020-12-24")
u = day("2023-01-20")
q("date_between", "sorted_from_and_to", f"{DB}:49-53", f"(date_between sorted_date {d} {d})", rows=[0])
q("date_between", "sorted_to_only", f"{DB}:55-59", f"(date_between sorted_date _ {d})", rows=[0])
q("date_between", "sorted_from_only", f"{DB}:61-65", f"(date_between sorted_date {d} _)", rows=[0])
q("date_between", "unsorted_from_and_to", f"{DB}:67-71", f"(date_between unsorted_date {u} {u})", rows=[0])
q("date_between", "unsorted_to_only", f"{DB}:73-77", f"(date_between unsorted_date _ {u})", rows=[0])
q("date_between", "unsorted_from_only", f"{DB}:79-83", f"(date_between unsorted_date {u} _)", rows=[0])
q("date_between", "column_not_in_database", f"{DB}:85-90", "(date_between something_not_in_database 10957 _)",
  error="query")
q("date_between", "non_date_column", f"{DB}:92-96", "(date_between primaryKey 18262 _)", error="query")

LF = f"{SEX}/lineage_filter.test.cpp"
q("lineage_filter", "lineage", f"{LF}:88-95", '(lineage pango_lineage "BASE.1" 0)', rows=[0, 1])
q("lineage_filter", "including_sublineages", f"{LF}:97-107", '(lineage pango_lineage "BASE.1" 1)', rows=[0, 1, 2])
q("lineage_filter", "null", f"{LF}:109-113", "(lineage pango_lineage null 0)", rows=[3])
q("lineage_filter", "null_including_sublineages", f"{LF}:115-121", "(lineage pango_lineage null 1)", rows=[3])
q("lineage_filter", "explicit_do_not_follow", f"{LF}:159-170", '(lineage pango_lineage "BASE.1" 1)', rows=[0, 1, 2])

OR = f"{SEX}/or.test.cpp"
SW, GE, FR, USA = ('(str_eq country "Switzerland")', '(str_eq country "Germany")', '(str_eq country "France")',
                   '(str_eq country "USA")')
q("or", "nested_or_same_column", f"{OR}:463-477", f"(or (or {SW} {GE}) {FR})", rows=[0, 1, 3, 4, 5])
q("or", "deeply_nested_or", f"{OR}:479-492", f"(or {SW} {GE})", rows=[0, 1, 3, 5])
q("or", "single_child_unwrapped", f"{OR}:494-500", f"(or {SW})", rows=[0, 3])
q("or", "string_equals_merged", f"{OR}:502-516", f"(or {SW} {GE} {FR})", rows=[0, 1, 3, 4, 5])
q("or", "mixed_columns", f"{OR}:518-532", f'(or {USA} (str_eq region "Europe"))', rows=[0, 1, 2, 3, 4, 5])
q("or", "with_and", f"{OR}:534-546", f'(or (and {SW} (str_eq region "Europe")) {USA})', rows=[0, 2, 3])
q("or", "empty_children", f"{OR}:548-552", "(false)", rows=[])

for name, lines, ref, rows, local, per_symbol in vsi_cases:
    for sym, want in per_symbol.items():
        q(f"vsi_{name}", f"{name}_{'gap' if sym == '-' else sym}", f"{VSI}:{lines}", f"(sym_eq s 0 {sym})", rows=want)

for name, lines, rows, expect in hci_cases:
    for pos, want in expect.items():
        q(f"hci_{name}", f"{name}_at_{pos}", f"{HCI}:{lines}", f"(covered s {pos})", rows=want)

BA = "src/rhydb/query_engine/operators/bitmap_aggregation_node.test.cpp"
aggregate_cases = [
    {"dataset": "bitmap_aggregation", "name": "two_positions", "ref": f"{BA}:72-82", "filter": "(true)",
     "dims": ["segment1@0", "segment1@1"], "expect": [["A", "T", 2], ["C", "A", 1], ["N", "N", 1]]},
    {"dataset": "bitmap_aggregation", "name": "with_filter", "ref": f"{BA}:84-93", "filter": "(has_mut segment1 0)",
     "dims": ["segment1@0", "segment1@1"], "expect": [["C", "A", 1]]},
    {"dataset": "bitmap_aggregation", "name": "amino_acid", "ref": f"{BA}:95-100", "filter": "(true)",
     "dims": ["gene1@1"], "expect": [["*", 4]]},
    {"dataset": "bitmap_aggregation", "name": "position_out_of_range", "ref": f"{BA}:129-134", "filter": "(true)",
     "dims": ["segment1@5"], "expect_error": "query"},
    {"dataset": "bitmap_aggregation", "name": "indexed_column", "ref": f"{BA}:139-147", "filter": "(true)",
     "dims": ["region"], "expect": [["Asia", 1], ["Europe", 3]]},
    {"dataset": "bitmap_aggregation", "name": "sequence_and_indexed_column", "ref": f"{BA}:153-162",
     "filter": "(true)", "dims": ["segment1@0", "region"],
     "expect": [["A", "Europe", 2], ["C", "Europe", 1], ["N", "Asia", 1]]},
    {"dataset": "bitmap_aggregation_nulls", "name": "null_two_positions", "ref": f"{BA}:204-216", "filter": "(true)",
     "dims": ["segment1@0", "segment1@1"], "expect": [["A", "T", 2], ["C", "A", 1], [None, None, 1]]},
    {"dataset": "bitmap_aggregation_nulls", "name": "null_amino_acid", "ref": f"{BA}:218-226", "filter": "(true)",
     "dims": ["gene1@0"], "expect": [["M", 3], [None, 1]]},
    {"dataset": "bitmap_aggregation_nulls", "name": "null_mixed_positions", "ref": f"{BA}:232-246",
     "filter": "(true)", "dims": ["segment1@0", "gene1@0"],
     "expect": [["A", "M", 2], ["C", None, 1], [None, "M", 1]]},
    {"dataset": "bitmap_aggregation_ambiguous", "name": "ambiguous_codes", "ref": f"{BA}:272-281",
     "filter": "(true)", "dims": ["segment1@0"], "expect": [["A", 1], ["R", 2], ["Y", 1]]},
]

mutation_cases = [
    {"dataset": "union_all", "name": "mutations_country_CH", "ref": "src/rhydb/query_engine/operators/"
     "union_all_node.test.cpp:185-194", "filter": '(str_eq country "CH")', "column": "segment1",
     "min_proportion": 0.0, "expect": [[1, "A", "T", 2, 2, 1.0]]},
    {"dataset": "union_all", "name": "mutations_country_DE", "ref": "src/rhydb/query_engine/operators/"
     "union_all_node.test.cpp:185-194", "filter": '(str_eq country "DE")', "column": "segment1",
     "min_proportion": 0.0, "expect": [[1, "A", "T", 2, 2, 1.0]]},
]

local_reference_cases = [
    {"dataset": f"vsi_{name}", "column": "s", "ref": f"{VSI}:{lines}", "expect": local}
    for name, lines, ref, rows, local, per_symbol in vsi_cases] + [
    {"dataset": "adapts_local_reference", "column": "test_column", "ref": "src/rhydb/storage/column/"
     "sequence_column.test.cpp:159-189", "expect": "ACTT"},
    {"dataset": "finalize_adapts", "column": "test_column", "ref": "src/rhydb/storage/column/"
     "sequence_column.test.cpp:191-212", "expect": "AAGT"},
    {"dataset": "symbol_equals", "column": "segment1", "ref": "derived: position 5 has reference N and one covered "
     "row with T", "expect": "ATGCT"},
]

with open(os.path.join(HERE, "reference_kats.json"), "w") as f:
    json.dump({"operator_cases": operator_cases, "datasets": datasets, "query_cases": query_cases,
               "mutation_cases": mutation_cases, "aggregate_cases": aggregate_cases,
               "local_reference_cases": local_reference_cases}, f, indent=None, separators=(",", ":"))
print(f"{len(operator_cases)} operator cases, {len(query_cases)} query cases, {len(mutation_cases)} mutation cases, "
      f"{len(aggregate_cases)} aggregate cases, {len(local_reference_cases)} local-reference cases")
