"""Loads the reference known-answer vectors (tests/golden/reference_kats.json) into TableSpecs."""
import json
import os

import numpy as np

from silo_amd import table

KATS = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_kats.json")))


def dataset_spec(name: str) -> table.TableSpec:
    """Keys of a dataset: "chunks" (row layout); "seq" {column: [alphabet, reference, rows]} with
    aligned rows (None = null); "placed" {column: [alphabet, reference, rows]} with rows
    [sequence, offset] placed in a genome of the reference's length, positions outside the sequence
    missing; "dict" {column: values} (None = null; ids in sorted name order, plus "extra_values"
    with no rows, and "lineage_parents" {child: parent} building a lineage index); "date" / "int"
    {column: values}; "bitmaps": literal row sets (row indices of the one-chunk layouts used)."""
    d = KATS["datasets"][name]
    spec = table.TableSpec(chunk_sizes=np.asarray(d["chunks"], np.uint32))
    n = spec.n_rows
    for col, (alphabet, ref, rows) in d.get("seq", {}).items():
        spec.seq_columns.append(table.seq_column_from_strings(col, ref, rows, alphabet))
    missing = {0: "N", 1: "X"}
    for col, (alphabet, ref, rows) in d.get("placed", {}).items():
        L = len(ref)
        aligned = [None if r is None else (missing[alphabet] * r[1] + r[0] + missing[alphabet] * L)[:L] for r in rows]
        spec.seq_columns.append(table.seq_column_from_strings(col, ref, aligned, alphabet))
    for col, values in d.get("dict", {}).items():
        names = sorted(set(v for v in values if v is not None) | set(d.get("extra_values", {}).get(col, [])))
        ids = np.array([names.index(v) if v is not None else 0 for v in values], np.uint32)
        nulls = np.array([v is None for v in values], np.uint8)
        parents = d.get("lineage_parents", {}).get(col)
        parent = None
        if parents is not None:
            parent = np.array([names.index(parents[v]) if v in parents else -1 for v in names], np.int32)
        spec.dict_columns.append(table.DictColumnData(col, ids, nulls, names, parent))
    for key, is_date in (("date", True), ("int", False)):
        for col, values in d.get(key, {}).items():
            spec.int_columns.append(table.IntColumnData(col, is_date, np.asarray(values, np.int32), np.zeros(n, np.uint8)))
    spec.bitmaps = [np.asarray(b, np.uint32) for b in d.get("bitmaps", [])]
    return spec


def aggregate_names(spec, dims, groups):
    """Groups (ids per dimension, count) as [value or symbol text or None, ..., count]."""
    out = []
    for key, count in groups:
        row = []
        for dim, v in zip(dims, key):
            if v is None:
                row.append(None)
            elif "@" in dim:
                col = next(c for c in spec.seq_columns if c.name == dim.split("@")[0])
                row.append(symbol_chars(col.alphabet)[v])
            else:
                row.append(next(c for c in spec.dict_columns if c.name == dim).names[v])
        out.append(row + [count])
    return out


def operator_spec(case) -> table.TableSpec:
    spec = table.TableSpec(chunk_sizes=np.asarray(case["chunks"], np.uint32))
    spec.bitmaps = [np.asarray(b, np.uint32) for b in case["bitmaps"]]
    return spec


def row_ids(spec, indices):
    g = spec.global_ids()
    return [int(g[i]) for i in indices]


def symbol_chars(alphabet: int) -> str:
    return table.alphabet_chars(alphabet)
